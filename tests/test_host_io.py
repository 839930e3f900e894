"""CPU tests of the host-side inputs and outputs: params.in reader, system.in
reader, RANMAR and setvel (against the oracle's restatement, which is pinned to
the reference's shipped bcc_iron test), checkpoint records."""
import numpy as np
import pytest

from moldy_amd import initial, moldin
from moldy_amd.params import FS_TO_TIME, GPA_TO_PRESS, ParamsError, parse_params

PARAMS = """#
# key = value pairs; comments and blank lines are skipped

title1 = "Input to a test run"   # a quoted title
title2 = "second title"
file_system          = "25^3.in"   # trailing comment
file_textout = output66.txt     # unquoted string
boxmass = 50.        # per particle
deltat  = 1.0d0
rpad    = 1.0D0
temprq  = 600.d0
press   = 0.1
ivol    = 0
iquen   = 0
straintensor      = 1.0 0.0 0.0 0.0 1.0 0.0 0.0 0.0 1.0
straintensor_row2 = 0.0 1.010 0.0
nm      = 31250
nsteps    = 10
nbrupdate = 2000
restart   = 0
nprint    = 10
nchkpt    = 1000
nnbrs     = 150
iverlet   = 1
dumpx1  = .False.
UseLookup = .true.
"""


def test_params_reader():
    p = parse_params(PARAMS)
    assert p.title1 == "Input to a test run" and p.title2 == "second title"
    assert p.file_system == "25^3.in" and p.file_textout == "output66.txt"
    assert p.deltat == 1.0 * FS_TO_TIME and p.deltat_fs == 1.0
    assert p.press == 0.1 * GPA_TO_PRESS
    assert p.rpad == 1.0 and p.temprq == 600.0
    assert (p.ivol, p.iquen, p.nm, p.nsteps, p.ntcm, p.nprint, p.nchkpt, p.nnbrs, p.iverlet) == \
        (0, 0, 31250, 10, 2000, 10, 1000, 150, 1)
    assert p.uselookup is True and p.dumpx1 is False
    assert p.bmass == 31250 * 50.0
    assert p.bdel2 == p.deltat * p.deltat / (2.0 * p.bmass)
    np.testing.assert_array_equal(p.strx, [[1, 0, 0], [0, 1.01, 0], [0, 0, 1]])
    assert parse_params("").ivol == -1 and parse_params("").iverlet == 0   # type defaults
    with pytest.raises(ParamsError):
        parse_params("nosuchkey = 3\n")
    with pytest.raises(ParamsError):
        parse_params("nsteps 10\n")


SYSTEM_IN = """2\t\t# Number of particles included in this file
3 2 4\t# Number of replications of the system in x,y,z
2.870 0.0 0.0 # Size of the system volume (3x3)
0.0 2.870 0.0 #
0.0 0.0 2.870 #
0.0 0.0 0.0 26 55.847
0.5 0.5 0.5 26 55.847
"""


def test_system_reader_and_setvel_match_oracle(oracle):
    rng = initial.Ranmar()
    b0, s, z, mass = initial.read_system(SYSTEM_IN, rng, dsp=0.07)
    ob0, os_, ox1, oz, omass = oracle.moldin_new_run([[0, 0, 0], [0.5, 0.5, 0.5]], [26, 26], [55.847, 55.847],
                                                     [3, 2, 4], np.diag([2.87] * 3), 0.07, 600.0, 1.0)
    np.testing.assert_array_equal(b0, ob0)
    np.testing.assert_array_equal(s, os_)
    np.testing.assert_array_equal(z, oz)
    x1 = initial.setvel(rng, z, mass, b0, 1.0 * FS_TO_TIME, initial.rqke_init(600.0, len(z)))
    np.testing.assert_array_equal(x1, ox1)
    assert np.abs(x1.sum(0)).max() < 1e-14        # centre-of-mass drift removed
    # kinetic energy equals rqke (setvel's scaling)
    v = x1 @ b0.T
    ke = 0.5 * np.sum(mass[:, None] * v * v) / (1.0 * FS_TO_TIME) ** 2
    assert abs(ke - initial.rqke_init(600.0, len(z))) < 1e-12 * ke


def test_ranmar_matches_oracle(oracle):
    a = initial.Ranmar(12, 34, 56, 78)
    b = oracle.Ranmar(12, 34, 56, 78)
    np.testing.assert_array_equal([a.rand1() for _ in range(5000)], b.rand1(5000))


def test_checkpoint_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 17
    arrs = [rng.normal(size=(n, 3)) for _ in range(4)]
    bs = [rng.normal(size=(3, 3)) for _ in range(4)]
    z = np.full(n, 26, dtype=np.int32)
    mass = np.full(n, 55.847)
    thm = rng.normal(size=24)
    path = str(tmp_path / "chk.dat")
    counters = dict(nlcx=3, nlcy=4, nlcz=5, prevsteps=0, currentstep=10, laststep=10, lastprint=10, lastchkpt=10,
                    ntc=7, iverlet=1, ivol=0)
    scal = dict(deltat=0.1, press=0.2, temprq=300.0, bmass=850.0, bdel2=1e-5)
    moldin.write_checkpoint(path, counters, scal, *arrs, bs, z, mass, np.ones(n, dtype=np.int32),
                            np.array([26], dtype=np.int32), thm, 10)
    ck = moldin.read_checkpoint(path, n)
    for k in range(4):
        np.testing.assert_array_equal(ck["x%d" % k], arrs[k])
        np.testing.assert_array_equal(ck["b%d" % k], bs[k])
    np.testing.assert_array_equal(ck["thm"], thm)
    assert ck["numsum"] == 10 and ck["counters"] == counters and ck["scalars"] == scal
    # record lengths as write_checkpoint_file lays them out: 3n+9 doubles per state record
    raw = open(path, "rb").read()
    first = int(np.frombuffer(raw, "<i4", 1, 0)[0])
    second = int(np.frombuffer(raw, "<i4", 1, first + 8)[0])
    assert second == 8 * (3 * n + 9)


class _BoxOnly:
    """Holds b1 and b2 the way the engine and the oracle expose them."""

    def __init__(self, b1, b2):
        self.t = {"b1": np.array(b1, dtype=np.float64), "b2": np.array(b2, dtype=np.float64)}

    def tensor(self, name):
        return self.t[name].copy()

    def set_tensor(self, name, v):
        self.t[name] = np.array(v, dtype=np.float64)


def test_box_quench_known_answer(oracle):
    """iquen = 2 (src/moldin.F90:845-853): b1(i,j) survives only when it has
    the sign of b2(i,j); opposite signs and a zero b2 zero it."""
    b1 = [[1e-3, -2e-4, 3e-5], [-4e-6, 0.0, 5e-4], [7e-5, -8e-5, -9e-4]]
    b2 = [[2e-9, 1e-9, -3e-9], [-1e-9, 4e-9, 0.0], [7e-10, -6e-10, -1e-12]]
    want = [[1e-3, 0.0, 0.0], [-4e-6, 0.0, 0.0], [7e-5, -8e-5, -9e-4]]
    box = _BoxOnly(b1, b2)
    moldin._box_quench(box)
    np.testing.assert_array_equal(box.t["b1"], want)
    np.testing.assert_array_equal(box.t["b2"], b2)
    box = _BoxOnly(b1, b2)
    oracle.OracleSystem.box_quench(box)
    np.testing.assert_array_equal(box.t["b1"], want)
