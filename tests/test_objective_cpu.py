"""The tuning objective's oracle (oracle/objective.py) pinned on the CPU.

``metrics.compute_2d_objective`` cannot run here (scikit-image is absent), so
the oracle restates it with the same numpy / scipy.ndimage calls and its own
Zhang-Suen thinning.  Pins:

* thinning: the algorithm's invariants (result inside the input, idempotent,
  one-pixel lines and end points kept, thick bars reduced to one pixel with
  their 8-connected components preserved) and a scalar implementation that
  visits pixels one by one; agreement with scikit-image itself is parity
  unpinned;
* the 7 x 7 median with scipy's "reflect" border: a brute-force median of the
  mirrored windows;
* the objective: closed forms on a straight shock (I_curve = 0, O = 0, TV,
  E_perp = u^2 |S|, E0 = v^2 |band|), which a tangent swapped for the
  normal, a wrong band or a wrong normalisation would miss.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import objective as ob


def _scalar_zhang_suen(img):
    """Zhang & Suen (1984), pixel by pixel: each subiteration marks from the
    unchanged image, then deletes."""
    a = np.pad(np.asarray(img, bool), 1).astype(int)
    m, n = a.shape
    while True:
        changed = False
        for step in (0, 1):
            marks = []
            for i in range(1, m - 1):
                for j in range(1, n - 1):
                    if not a[i, j]:
                        continue
                    P = [a[i - 1, j], a[i - 1, j + 1], a[i, j + 1], a[i + 1, j + 1],
                         a[i + 1, j], a[i + 1, j - 1], a[i, j - 1], a[i - 1, j - 1]]
                    B = sum(P)
                    A = sum(1 for k in range(8) if P[k] == 0 and P[(k + 1) % 8] == 1)
                    p2, p4, p6, p8 = P[0], P[2], P[4], P[6]
                    if step == 0:
                        ok = p2 * p4 * p6 == 0 and p4 * p6 * p8 == 0
                    else:
                        ok = p2 * p4 * p8 == 0 and p2 * p6 * p8 == 0
                    if 2 <= B <= 6 and A == 1 and ok:
                        marks.append((i, j))
            for i, j in marks:
                a[i, j] = 0
            changed = changed or bool(marks)
        if not changed:
            return a[1:-1, 1:-1].astype(bool)


def _blobs(seed, shape=(40, 36), density=0.5):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.random(shape), 2.0)
    return f > np.quantile(f, 1 - density)


def _components(img):
    return ndimage.label(img, structure=np.ones((3, 3)))[1]


@pytest.mark.parametrize("seed", range(6))
def test_thinning_matches_scalar_and_invariants(seed):
    img = _blobs(seed)
    sk = ob.zhang_suen(img)
    np.testing.assert_array_equal(sk, _scalar_zhang_suen(img))
    assert not (sk & ~img).any()                                 # inside the input
    np.testing.assert_array_equal(ob.zhang_suen(sk), sk)          # idempotent
    win = sk[:-1, :-1] & sk[1:, :-1] & sk[:-1, 1:] & sk[1:, 1:]
    assert win.sum() <= 2                                         # (almost) no 2 x 2 blocks left
    assert sk.sum() < img.sum() / 2


def test_thinning_keeps_lines_and_end_points():
    img = np.zeros((20, 20), bool)
    img[3, 2:15] = True                       # horizontal
    img[6:18, 17] = True                      # vertical
    for k in range(8):
        img[8 + k, 2 + k] = True              # diagonal
    img[12:19, 12] = True                     # L shape
    img[18, 12:16] = True
    np.testing.assert_array_equal(ob.zhang_suen(img), img)


@pytest.mark.parametrize("w,h", [(3, 20), (5, 31), (9, 40), (7, 7)])
def test_thick_bars_thin_to_connected_curves(w, h):
    img = np.zeros((h + 6, w + 12), bool)
    img[3:3 + h, 6:6 + w] = True
    sk = ob.zhang_suen(img)
    assert _components(sk) == 1
    # one pixel wide across the bar
    assert (sk.sum(axis=1) <= 2).all()
    assert sk[3 + h // 2, 6:6 + w].sum() == 1
    # two separate bars stay two components
    img2 = img.copy()
    img2[:, -4:-1] = True
    assert _components(ob.zhang_suen(img2)) == 2


def test_median_filter_against_brute_force():
    rng = np.random.default_rng(3)
    for shape in ((9, 11), (4, 13), (2, 5), (16, 16)):
        r = rng.random(shape)
        r[r > 0.7] = 0.5                        # ties
        got = ob.objective_fields(r)[3]
        m, n = shape

        def refl(k, size):   # d c b a | a b c d | d c b a, repeated
            k %= 2 * size
            return k if k < size else 2 * size - 1 - k

        want = np.empty(shape)
        for i in range(m):
            for j in range(n):
                rows = [refl(i + d, m) for d in range(-3, 4)]
                cols = [refl(j + d, n) for d in range(-3, 4)]
                want[i, j] = np.sort(r[np.ix_(rows, cols)].ravel())[24]
        np.testing.assert_array_equal(got, want)


def test_objective_on_a_straight_shock():
    """rho steps from 1 to 3 across column 20 of a 32 x 48 grid; u = 0.7,
    v = 0.4 everywhere."""
    nx, ny, c = 32, 48, 20
    rho = np.ones((nx, ny))
    rho[:, c:] = 3.0
    u = np.full((nx, ny), 0.7)
    v = np.full((nx, ny), 0.4)
    J, m = ob.compute_2d_objective(rho, u, v, None)
    S, C, band, med = ob.objective_fields(rho)
    # the gradient lives on columns c-1 and c; the band is 2 columns wider each side
    assert S.sum() == 2 * nx and S[:, c - 1].all() and S[:, c].all()
    assert band.sum() == 6 * nx and band[:, c - 3:c + 3].all()
    # the skeleton is a straight column: zero residual (lstsq rounding aside)
    assert C.sum() >= 10 and len(set(np.nonzero(C)[1])) == 1
    assert m["I_curve"] <= 1e-12
    assert m["O"] == 0.0                            # the median keeps a step
    assert m["TV"] == nx * 2.0 / (nx * ny)
    # where the gradient is nonzero (the two columns of S) the level-set
    # tangent is along x and vperp = -u; elsewhere the normal is 0
    assert m["E_perp"] == pytest.approx(0.7 ** 2 * S.sum(), rel=1e-15)
    # E0 = median(|v| over the band)^2 |band|
    assert m["cE"] == pytest.approx(0.7 ** 2 * S.sum() / (0.4 ** 2 * band.sum()), rel=1e-15)
    tv0 = np.mean(rho)
    assert m["cTV"] == pytest.approx(m["TV"] / tv0, rel=1e-15)
    assert J == pytest.approx(m["cI"] + m["cE"] + 50 * m["cO"] + m["cTV"], rel=1e-15)


def test_objective_without_shock_and_weights():
    rho = np.full((12, 10), 2.0)
    u = np.zeros((12, 10))
    v = np.ones((12, 10))
    J, m = ob.compute_2d_objective(rho, u, v, None, weights={"I": 1, "E": 1, "O": 1, "TV": 1})
    # G = 0 everywhere, so S, band and the skeleton are empty
    assert m["I_curve"] == 0 and m["E_perp"] == 0 and m["O"] == 0 and m["TV"] == 0
    assert J == 0.0


def test_fallback_search_finds_the_minimum_of_a_known_function():
    """The built-in Bayesian search (used when scikit-optimize is absent) on
    the reference's box and constraint: a smooth bowl with its minimum at
    (0.3, 4.0) and the 1e6 penalty for jp_lo >= jp_hi; 30 evaluations get
    within 0.02 of the minimum value, deterministic for a seed."""
    from wenox import tuning

    def f(x):
        if x[0] >= x[1]:
            return tuning.PENALTY
        return (x[0] - 0.3) ** 2 + 0.01 * (x[1] - 4.0) ** 2

    x1, f1, xs, fs = tuning.gp_ei_minimize(f, tuning.SPACE, 30, 8, 0)
    x2, f2, _, _ = tuning.gp_ei_minimize(f, tuning.SPACE, 30, 8, 0)
    assert len(xs) == 30 and f1 == min(fs)
    assert f1 == f2 and np.array_equal(x1, x2)
    assert f1 < 0.02
    for (lo, hi), v in zip(tuning.SPACE, x1):
        assert lo <= v <= hi
