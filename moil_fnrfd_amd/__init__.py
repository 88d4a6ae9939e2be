"""MI355X-native MOIL energy/force evaluation (eforce hot path)."""
