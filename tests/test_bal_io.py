"""BAL text reader/writer (examples/bal/model.rs:59-113) in the native library: the
reference's own parser cases (tests/bal.rs:43-61), the token grammar of Rust's `str::parse`
that the reference relies on, bit-exact write/read round trips, and the independent BAL
objective against the oracle's cost. Host code only: runs without a GPU."""
import numpy as np
import pytest

from fagra_amd import bal, synthetic
from oracle import oracle

VALID = "1 1 1 0 0 0 0 0 0 0 0 0 0 100 0 0 0 0 -5"


def test_reference_parser_cases():
    # tests/bal.rs:43-61
    p = bal.parse(VALID)
    assert p.n_cameras == 1 and p.points[0, 2] == -5.0
    assert p.cameras[0, 6] == 100.0 and p.n_observations == 1
    cases = {
        "": "truncated BAL input",
        "0 1 1": "empty BAL problem",
        "1 1 1 1 0 0 0": "invalid observation index or value",
        "1 1 1 0 1 0 0": "invalid observation index or value",
        "1 1 1 0 0 NaN 0": "invalid observation index or value",
        VALID + " 0": "trailing BAL tokens",
        VALID.replace("100", "inf"): "nonfinite parameter",
        VALID.rsplit(" ", 1)[0]: "truncated BAL input",
    }
    for text, msg in cases.items():
        with pytest.raises(bal.BalError) as e:
            bal.parse(text)
        assert str(e.value) == msg, text


@pytest.mark.parametrize("tok,value", [("+1", None), ("1.", 1.0), (".5", 0.5), ("1e3", 1e3), ("1E+3", 1e3),
                                       ("-0", -0.0), ("2.5e-3", 2.5e-3), ("1e-400", 0.0),
                                       ("4.9406564584124654e-324", 5e-324), ("-7.25", -7.25)])
def test_float_tokens_rust_accepts(tok, value):
    if value is None:  # an index token with '+'
        p = bal.parse("1 1 1 +0 +0 0 0 " + " ".join(["1"] * 12))
        assert p.obs_camera[0] == 0
        return
    p = bal.parse(f"1 1 1 0 0 {tok} 0 " + " ".join(["1"] * 12))
    assert p.obs_pixel[0, 0] == value
    assert np.signbit(p.obs_pixel[0, 0]) == np.signbit(value)


@pytest.mark.parametrize("text,msg", [
    ("1 1 1 0 0 0x10 0", "invalid BAL token"),
    ("1 1 1 0 0 1e 0", "invalid BAL token"),
    ("1 1 1 0 0 . 0", "invalid BAL token"),
    ("1 1 1 0 0 1_0 0", "invalid BAL token"),
    ("1 1 1 -1 0 0 0", "invalid BAL token"),
    ("1 1 1 0.0 0 0 0", "invalid BAL token"),
    ("18446744073709551616 1 1", "invalid BAL token"),
    ("1 1 1 0 0 1e400 0", "invalid observation index or value"),
    ("1 1 1 0 0 -Infinity 0", "invalid observation index or value"),
    ("1 1", "truncated BAL input"),
    ("1 1 1 0 0 0", "truncated BAL input"),
    # counts far beyond the text end as truncated input, with nothing allocated for them
    ("1000000000000 1 1 0 0 0 0", "truncated BAL input"),
])
def test_rejected_tokens(text, msg):
    with pytest.raises(bal.BalError) as e:
        bal.parse(text)
    assert str(e.value) == msg


def test_whitespace_kinds():
    text = VALID.replace(" ", "\t", 3).replace(" ", "\r\n", 2).replace(" ", "\x0b\x0c  ", 1) + "\n\n"
    p = bal.parse(text)
    q = bal.parse(VALID)
    assert np.array_equal(p.cameras, q.cameras) and np.array_equal(p.points, q.points)


def test_round_trip_is_bit_exact():
    p = synthetic.generate("ring16", seed=4, n_cameras=6, n_points=150)
    q = bal.parse(bal.format(p))
    for a in ("cameras", "points", "obs_camera", "obs_point", "obs_pixel"):
        x, y = getattr(p, a), getattr(q, a)
        assert x.dtype == y.dtype and x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8)), a
    # awkward values: subnormals, signed zeros, extreme exponents, shortest forms
    rng = np.random.default_rng(0)
    cams = rng.standard_normal((2, 9)) * 10.0 ** rng.integers(-300, 300, (2, 9))
    cams[0, :4] = [5e-324, -0.0, 1.7976931348623157e308, 0.1]
    pts = np.array([[2.2250738585072014e-308, -1e-320, 1.0 / 3.0]])
    p2 = synthetic.BalProblem(name="x", cameras=cams, points=pts, obs_camera=np.array([1, 0]),
                              obs_point=np.array([0, 0]), obs_pixel=np.array([[0.5, -2.0], [1e300, -1e-300]]))
    q2 = bal.parse(bal.format(p2))
    assert np.array_equal(q2.cameras.view(np.uint64), cams.view(np.uint64))
    assert np.array_equal(q2.points.view(np.uint64), pts.view(np.uint64))
    assert np.array_equal(q2.obs_pixel.view(np.uint64), p2.obs_pixel.view(np.uint64))
    # replacement states and the file helpers
    text = bal.format(p, cameras=p.cameras + 1.0)
    assert np.array_equal(bal.parse(text).cameras, p.cameras + 1.0)


def test_writer_rejects_what_the_reader_would():
    p = synthetic.generate("ring16", seed=4, n_cameras=3, n_points=20)
    bad = p.cameras.copy()
    bad[0, 0] = np.nan
    with pytest.raises(ValueError):
        bal.format(p, cameras=bad)
    with pytest.raises(ValueError):
        bal.format(p, cameras=p.cameras[:2])


def test_file_helpers(tmp_path):
    p = synthetic.generate("ring16", seed=5, n_cameras=3, n_points=20)
    path = tmp_path / "problem.txt"
    bal.write(str(path), p)
    q = bal.read(str(path))
    assert q.name == "problem.txt" and np.array_equal(q.points, p.points)
    assert bal.identify(str(path)) is None
    assert len(bal.LADYBUG_SHA256) == 3


@pytest.mark.parametrize("seed", [0, 3])
def test_objective_matches_oracle_cost(seed):
    # the runner's independent objective (examples/bal.rs:202-219) against the oracle's
    # evaluator cost (examples/bal/model.rs:201-237) on the same state
    p = synthetic.generate("ring16", seed=seed, n_cameras=5, n_points=120)
    st, c = oracle.BA(p, huber=0.0).cost()
    assert st == oracle.OK
    assert abs(bal.objective(p, p.cameras, p.points) - c) <= 1e-12 * c
    # and a hand-evaluated pixel (tests/bal.rs:106-111): f=100, k1=0.1, k2=0.01, X=(1,2,-2)
    one = synthetic.BalProblem(name="x", cameras=np.array([[0, 0, 0, 0, 0, 0, 100.0, 0.1, 0.01]]),
                               points=np.array([[1.0, 2.0, -2.0]]), obs_camera=np.array([0]),
                               obs_point=np.array([0]), obs_pixel=np.zeros((1, 2)))
    px = 50.0 * (1 + 0.1 * 1.25 + 0.01 * 1.25 * 1.25)
    assert abs(bal.objective(one, one.cameras, one.points) - 0.5 * (px ** 2 + (2 * px) ** 2)) < 1e-9
