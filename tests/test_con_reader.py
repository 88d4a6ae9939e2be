"""The library's .con reader (moil_con_read, MOIL's rconn semantics,
generic/rconn.f:46-497) against the topology the file was written from:
every array identical, for water boxes, heteropolymers in water, chains
without water and a single water.  The written files are the ones MOIL's
own rconn reads for the golden fixtures (the fixture topology checksums
pin those), so the reader hands moil_gpu_init exactly what MOIL reads.
Host only: no GPU call."""
import os

import numpy as np
import pytest

from moil_fnrfd_amd import abi, configs
from moil_fnrfd_amd.systems import monomer_flags, write_con
from tests._cases import make_case


def _read(path):
    from moil_fnrfd_amd.eforce import MoilCon
    return MoilCon(path)


CASES = ['water216', 'hetero_small', 'chains_only', 'chains_one_water', 'hetero_small_out',
         'hetero_small_arith']


@pytest.mark.parametrize('name', CASES)
def test_reader_returns_the_written_topology(name, tmp_path):
    s, cut, ew, sqz = make_case(name)
    t = s.topo
    path = os.path.join(tmp_path, name + '.con')
    write_con(path, t)
    c = _read(path)
    T = c.topo
    assert (T.npt, T.totmon) == (t.npt, t.totmon)
    assert (T.nb, T.nangl, T.ntors, T.nimp, T.totspe, T.nwaters) == (
        len(t.ib1), len(t.iangl1), len(t.itor1), len(t.iimp1), len(t.spec1), t.nwaters)
    eq = np.testing.assert_array_equal
    eq(c.array('poipt', t.totmon + 1), t.poipt)
    eq(c.array('monoflag', t.totmon), monomer_flags(t.moname))
    for f in ('ptchg', 'epsgm6', 'epsgm12'):
        eq(c.array(f, t.npt), getattr(t, f))
    eq(c.array('flagchr', t.npt), np.asarray(t.flagchr).astype(np.int32))
    eq(c.array('exc1', t.npt + 1), t.exc1)
    eq(c.array('exc2', len(t.exc2)), t.exc2)
    for f in ('ib1', 'ib2', 'kbond', 'req'):
        eq(c.array(f, len(t.ib1)), getattr(t, f))
    for f in ('iangl1', 'iangl2', 'iangl3', 'kangl', 'angleq'):
        eq(c.array(f, len(t.iangl1)), getattr(t, f))
    for f in ('itor1', 'itor2', 'itor3', 'itor4', 'period', 'ktors1', 'ktors2', 'ktors3',
              'phase1', 'phase2', 'phase3'):
        eq(c.array(f, len(t.itor1)), getattr(t, f))
    for f in ('iimp1', 'iimp2', 'iimp3', 'iimp4', 'kimp', 'impeq'):
        eq(c.array(f, len(t.iimp1)), getattr(t, f))
    eq(c.array('spec1', len(t.spec1)), t.spec1)
    eq(c.array('spec2', len(t.spec2)), t.spec2)
    eq(c.array('p14', len(t.spec1), 3), t.p14)
    eq(c.array('idxtip3', t.nwaters), t.idxtip3)
    assert c.arith == t.arith
    c.close()


def test_exclusion_records_longer_than_one_line(tmp_path):
    """An atom with more than ten exclusions continues its partner list on
    the next record, which MOIL's format(1x,10(i7,1x)) reads from column 1
    (format reversion to the group); the reader does the same."""
    s, cut, ew = configs.water216()
    path = os.path.join(tmp_path, 'w.con')
    write_con(path, s.topo)
    lines = open(path).read().splitlines()
    # rewrite the first exclusion record of the file as one atom excluding
    # 12 partners: a first line of ten (1x,i7,1x) items and a continuation
    # of two (i7,1x) items starting in column 1
    k = next(i for i, l in enumerate(lines) if l.startswith('~ atom number, number of excl'))
    partners = list(range(2, 14))
    head = ' %7d %7d' % (1, len(partners))
    first = ' ' + ''.join('%7d ' % p for p in partners[:10])
    cont = ''.join('%7d ' % p for p in partners[10:])
    # drop the old records of atoms 1.. up to the charge flags, keep totex consistent
    end = next(i for i, l in enumerate(lines) if l.startswith('~ Charge flags'))
    new = lines[:k + 1] + [head, first, cont] + lines[end:]
    # header count totex (7th integer of the counts record)
    cnt = lines[3]
    vals = [int(cnt[1 + 8 * i:9 + 8 * i]) for i in range(14)]
    vals[6] = len(partners)
    new[3] = ' ' + ''.join('%8d' % v for v in vals) + cnt[1 + 8 * 14:]
    path2 = os.path.join(tmp_path, 'w2.con')
    with open(path2, 'w') as f:
        f.write('\n'.join(new) + '\n')
    c = _read(path2)
    exc1 = c.array('exc1', s.topo.npt + 1)
    exc2 = c.array('exc2', len(partners))
    assert exc1[0] == 0 and exc1[1] == 12 and (exc1[2:] == 12).all()
    np.testing.assert_array_equal(exc2, partners)
    c.close()


def test_reader_refuses_what_the_library_does_not_support(tmp_path):
    from moil_fnrfd_amd.eforce import MoilGPUError
    s, cut, ew = configs.water216()
    path = os.path.join(tmp_path, 'w.con')
    write_con(path, s.topo)
    lines = open(path).read().splitlines()
    cnt = lines[3]
    for idx, what in ((8, 'LES'), (10, 'Morse')):
        vals = [int(cnt[1 + 8 * i:9 + 8 * i]) for i in range(14)]
        vals[idx] = 1
        bad = list(lines)
        bad[3] = ' ' + ''.join('%8d' % v for v in vals) + cnt[1 + 8 * 14:]
        p2 = os.path.join(tmp_path, 'bad%d.con' % idx)
        with open(p2, 'w') as f:
            f.write('\n'.join(bad) + '\n')
        with pytest.raises(MoilGPUError, match=what):
            _read(p2)
    with pytest.raises(MoilGPUError, match='cannot open'):
        _read(os.path.join(tmp_path, 'missing.con'))
    p3 = os.path.join(tmp_path, 'short.con')
    with open(p3, 'w') as f:
        f.write('\n'.join(lines[:40]) + '\n')
    with pytest.raises(MoilGPUError, match='error %d' % abi.EINVAL):
        _read(p3)


def _write_crd(path, s, title=('* synthetic system', '* written for the reader test')):
    """A coordinate file in MOIL's getcrd layout (generic/getcrd.f:100)."""
    t = s.topo
    mono_of = np.searchsorted(t.poipt[1:], np.arange(1, t.npt + 1))
    lines = list(title) + ['%7d' % t.npt]
    for i in range(t.npt):
        m = int(mono_of[i])
        x, y, z = s.coor[i]
        lines.append('%7d%7d %-4s %-4s%10.5f%10.5f%10.5f %-4s %-4s%10.5f' % (
            i + 1, m + 1, t.moname[m][:4], t.ptnm[i][:4], x, y, z, 'SYNT', str(m + 1)[:4], 0.0))
    with open(path, 'w') as f:
        f.write('\n'.join(lines) + '\n')


def test_coordinate_reader(tmp_path):
    from moil_fnrfd_amd.eforce import MoilGPUError, read_crd
    s, cut, ew, sqz = make_case('hetero_small')
    path = os.path.join(tmp_path, 's.crd')
    _write_crd(path, s)
    x = read_crd(path, s.topo.npt)
    # f10.5 fields: the values as written, to the last bit
    ref = np.array([[float('%10.5f' % v) for v in row] for row in s.coor])
    np.testing.assert_array_equal(x, ref)
    assert np.abs(x - s.coor).max() <= 5e-6
    lines = open(path).read().splitlines()
    bad = os.path.join(tmp_path, 'bad.crd')
    with open(bad, 'w') as f:
        f.write('\n'.join(lines[:3] + lines[4:]) + '\n')  # particle 1 missing
    with pytest.raises(MoilGPUError, match='missing'):
        read_crd(bad, s.topo.npt)
    with pytest.raises(MoilGPUError, match='does not match'):
        read_crd(path, s.topo.npt + 1)


def test_coordinate_reader_matches_moil_getcrd(tmp_path):
    """MOIL's own getcrd (through oracle/_ref/moil_dump) reads the same file
    to the same coordinates, bit for bit (the file's particle names are
    distinct within each monomer, where getcrd's name matching and the
    library's record order agree)."""
    from oracle import ref_runner
    if not ref_runner.available():
        pytest.skip('oracle/_ref/moil_dump not built')
    from moil_fnrfd_amd.eforce import read_crd
    s, cut, ew, sqz = make_case('hetero_small')
    path = os.path.join(tmp_path, 's.crd')
    _write_crd(path, s)
    d = ref_runner.run_moil(s, cut, ew, dump_lists=False, crd_file=path, timeout=300)
    np.testing.assert_array_equal(read_crd(path, s.topo.npt), d['COOR'].reshape(-1, 3))


def test_trajectory_writer_matches_moil_wdyncrd(tmp_path):
    """Three frames written by the library and by MOIL's own wdyncrd
    (oracle/_ref/moil_traj): the files are identical byte for byte."""
    import subprocess

    from moil_fnrfd_amd.eforce import MoilTraj
    from oracle import ref_runner
    exe = os.path.join(os.path.dirname(ref_runner.REF_EXE), 'moil_traj')
    s, cut, ew, sqz = make_case('hetero_small')
    npt = s.topo.npt
    rng = np.random.default_rng(9)
    frames = [s.coor + rng.normal(0.0, 0.1, s.coor.shape) for _ in range(3)]
    mine = os.path.join(tmp_path, 'lib.dcd')
    t = MoilTraj(mine, npt, len(frames))
    for x in frames:
        t.write(x)
    t.close()
    data = open(mine, 'rb').read()
    # layout: header records then 3 single-precision records per frame
    head = 4 + 84 + 4 + 4 + 84 + 4 + 4 + 4 + 4
    assert len(data) == head + len(frames) * 3 * (8 + 4 * npt)
    off = head + 4
    np.testing.assert_array_equal(np.frombuffer(data, np.float32, npt, off),
                                  frames[0][:, 0].astype(np.float32))
    if not os.path.exists(exe):
        pytest.skip('oracle/_ref/moil_traj not built')
    raw = os.path.join(tmp_path, 'frames.bin')
    np.ascontiguousarray(np.stack(frames), np.float64).tofile(raw)
    ref = os.path.join(tmp_path, 'moil.dcd')
    par = os.path.join(tmp_path, 'traj.par')
    with open(par, 'w') as f:
        f.write('%d %d\n%s\n%s\n' % (npt, len(frames), raw, ref))
    subprocess.run([exe, par], check=True, capture_output=True, timeout=60)
    assert open(ref, 'rb').read() == data
    # and MOIL's file reads back frame by frame, widened from single precision
    from moil_fnrfd_amd.eforce import MoilGPUError, read_traj
    for k, x in enumerate(frames):
        y, nf = read_traj(ref, npt, k)
        assert nf == len(frames)
        np.testing.assert_array_equal(y, x.astype(np.float32).astype(np.float64))
    with pytest.raises(MoilGPUError, match='not in the file'):
        read_traj(ref, npt, len(frames))
    with pytest.raises(MoilGPUError, match='does not match'):
        read_traj(ref, npt + 1, 0)


def test_force_norm_follows_moil():
    """moil_force_norm against force_norm's own sequential sum
    (pot/force_norm.f:20-25), bit for bit, on MOIL's gradient of a golden
    fixture; and against the closed form to rounding."""
    import math

    from moil_fnrfd_amd.eforce import force_norm
    from tests._util import load_golden
    dp = load_golden('hetero_small')['dpot']
    g = 0.0
    for x, y, z in dp:
        g = g + x * x
        g = g + y * y
        g = g + z * z
    ref = math.sqrt(g / (3 * len(dp)))
    assert force_norm(dp) == ref
    assert abs(ref - np.sqrt((dp ** 2).mean())) <= 1e-13 * ref
