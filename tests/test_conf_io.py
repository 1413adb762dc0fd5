"""BQCD configuration-file I/O (hb_conf_read / hb_conf_write, host only)
against files the reference itself wrote (conf_write with field_io and the
POSIX checksum; tests/golden/conf_l4446.npz, made by make_golden.py conf),
and the reverse: files this library writes read by the reference's conf_read
(oracle/_ref/bqcd_confio, when built).  No GPU is needed."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIO = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "bqcd_confio")
LIB = os.path.join(os.path.dirname(HERE), "bqcd-r399-quda_amd", "libbqcd_hip.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libbqcd_hip.so not built")


@pytest.fixture(scope="module")
def ref(golden):
    f = golden("conf_l4446")
    L = tuple(int(x) for x in f["L"])
    u = ol.Lattice(L).gauge_exp_random(int(f["seed"]), float(f["scale"]))
    files = {str(n): f["file_%d" % k].tobytes() for k, n in enumerate(f["names"])}
    return L, u, files


def unpack(files, d):
    for n, b in files.items():
        with open(os.path.join(d, n), "wb") as fh:
            fh.write(b)
    return os.path.join(d, [n for n in files if n.endswith(".info")][0])


def test_read_reference_files(ref):
    import bqcd_hip

    L, u, files = ref
    with tempfile.TemporaryDirectory() as d:
        info = unpack(files, d)
        r, plaq = bqcd_hip.conf_read(info, L)
    # the first two columns are the stored float64 values, the third is rebuilt;
    # the oracle's links are SU(3) to ~1e-14, which bounds the difference
    assert np.array_equal(r[:, :, :, :2], u[:, :, :, :2])
    assert np.max(np.abs(r - u)) < 5e-14
    lat = ol.Lattice(L)
    assert abs(plaq - (1.0 - lat.plaquette_sum(u) / (3 * 6 * 2 * lat.volh))) < 1e-12


def test_write_matches_reference_bytes(ref):
    import bqcd_hip

    L, u, files = ref
    with tempfile.TemporaryDirectory() as d:
        info = bqcd_hip.conf_write(L, u, d, run=1, traj=0, beta=5.29, kappa=0.1355, csw=1.9)
        for n, b in files.items():
            if n.endswith(".u"):
                assert open(os.path.join(d, n), "rb").read() == b, n
        ours = open(info).read().split(">BeginCheckSum")[1]
        theirs = files[os.path.basename(info)].decode().split(">BeginCheckSum")[1]
        assert ours.split() == theirs.split()   # file names, checksums, byte counts


@pytest.mark.skipif(not os.path.exists(CONFIO), reason="oracle/_ref/bqcd_confio not built")
def test_reference_reads_our_files():
    import bqcd_hip

    L = (4, 4, 6, 4)
    lat = ol.Lattice(L)
    u = lat.gauge_exp_random(17, 0.4)
    with tempfile.TemporaryDirectory() as d:
        info = bqcd_hip.conf_write(L, u, d, run=3, traj=12, beta=5.29, kappa=0.13, csw=1.0,
                                   plaq_energy=1.0 - lat.plaquette_sum(u) / (18 * 2 * lat.volh))
        ctl = "%d %d %d %d\nread\n%s\n%s\n" % (*L, d, info)
        p = subprocess.run([CONFIO], input=ctl.encode(), cwd=d, capture_output=True, timeout=60)
        assert p.returncode == 0, p.stdout.decode() + p.stderr.decode()
        r = np.fromfile(os.path.join(d, "u_read.bin"), "<f8").reshape(u.shape)
        ours, _ = bqcd_hip.conf_read(info, L)
    assert np.max(np.abs(r - u)) < 5e-14
    # the third column rebuilt here and by the reference's u_complete
    assert np.max(np.abs(ours - r)) < 1e-15


def test_corrupt_and_mismatched_files_are_rejected(ref):
    import bqcd_hip

    L, u, files = ref
    with tempfile.TemporaryDirectory() as d:
        info = unpack(files, d)
        with pytest.raises(bqcd_hip.HipError, match="lattice"):
            bqcd_hip.conf_read(info, (4, 4, 4, 8))
        slice3 = os.path.join(d, sorted(n for n in files if n.endswith(".u"))[3])
        b = bytearray(open(slice3, "rb").read())
        b[1000] ^= 0x10
        open(slice3, "wb").write(bytes(b))
        with pytest.raises(bqcd_hip.HipError, match="checksum"):
            bqcd_hip.conf_read(info, L)


def test_missing_truncated_and_unwritable_files_are_errors(ref):
    import bqcd_hip

    L, u, files = ref
    slices = sorted(n for n in files if n.endswith(".u"))
    with tempfile.TemporaryDirectory() as d:
        info = unpack(files, d)
        os.remove(os.path.join(d, slices[2]))
        with pytest.raises(bqcd_hip.HipError, match="cannot open"):
            bqcd_hip.conf_read(info, L)
        unpack(files, d)
        with open(os.path.join(d, slices[4]), "wb") as fh:
            fh.write(files[slices[4]][:100])
        with pytest.raises(bqcd_hip.HipError, match="wrong size"):
            bqcd_hip.conf_read(info, L)
        with pytest.raises(bqcd_hip.HipError, match="cannot write"):
            bqcd_hip.conf_write(L, u, os.path.join(d, "no", "such", "dir"), run=1, traj=0,
                                beta=5.29, kappa=0.13, csw=1.9)
