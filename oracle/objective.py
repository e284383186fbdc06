"""CPU oracle of the jp_cri tuning objective -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s CPU-baseline leg may import this module, as
the checker.

``compute_2d_objective`` restates ``metrics.compute_2d_objective``
(``metrics.py:1-77``) line by line with the same numpy and scipy.ndimage calls
(``median_filter`` size 7, mode "reflect"; ``binary_dilation`` with a 5 x 5
structure).  The reference thins the shock mask with
``skimage.morphology.skeletonize``, and scikit-image is not installed here:
``zhang_suen`` is a restatement of the published algorithm that function
implements for 2-D images (T. Y. Zhang and C. Y. Suen, "A fast parallel
algorithm for thinning digital patterns", CACM 27(3), 1984), with the image
padded by background.  Its agreement with scikit-image is parity unpinned; the
tests pin it by the algorithm's invariants instead.
"""
from __future__ import annotations

import numpy as np
from scipy.ndimage import binary_dilation, median_filter

# neighbour offsets P2..P9 (north, north-east, east, ..., north-west), axis 0
# pointing "south" as rows do
_NB = ((-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1), (-1, -1))


def zhang_suen(mask: np.ndarray) -> np.ndarray:
    """Zhang-Suen thinning of a 2-D boolean image; pixels outside are
    background.  Two parallel subiterations per pass, repeated until a pass
    removes nothing.  A pixel P1 is removed when 2 <= B(P1) <= 6 (foreground
    neighbours), A(P1) = 1 (0 -> 1 transitions around P2..P9,P2), and
    P2 P4 P6 = P4 P6 P8 = 0 (first subiteration) or P2 P4 P8 = P2 P6 P8 = 0
    (second)."""
    img = np.pad(np.asarray(mask, dtype=bool), 1).astype(np.uint8)
    m, n = img.shape
    while True:
        removed = False
        for first in (True, False):
            p = [img[1 + di:m - 1 + di, 1 + dj:n - 1 + dj] for di, dj in _NB]
            b = sum(x.astype(np.int32) for x in p)
            a = sum(((p[k] == 0) & (p[(k + 1) % 8] == 1)).astype(np.int32) for k in range(8))
            P2, P4, P6, P8 = p[0], p[2], p[4], p[6]
            if first:
                c = (P2 * P4 * P6 == 0) & (P4 * P6 * P8 == 0)
            else:
                c = (P2 * P4 * P8 == 0) & (P2 * P6 * P8 == 0)
            core = img[1:-1, 1:-1]
            kill = (core == 1) & (b >= 2) & (b <= 6) & (a == 1) & c
            if kill.any():
                core[kill] = 0
                removed = True
        if not removed:
            return img[1:-1, 1:-1].astype(bool)


def compute_2d_objective(rho, u, v, p, dx=1.0, dy=1.0, weights=None, skeletonize=zhang_suen):
    """metrics.py:1-77 (the objective of opt_jp_cro.py:41)."""
    # :6-12 shock detector
    rho_x = (np.pad(rho, ((1, 1), (0, 0)), "edge")[2:, :] - np.pad(rho, ((1, 1), (0, 0)), "edge")[:-2, :]) / (2 * dx)
    rho_y = (np.pad(rho, ((0, 0), (1, 1)), "edge")[:, 2:] - np.pad(rho, ((0, 0), (1, 1)), "edge")[:, :-2]) / (2 * dy)
    G = np.sqrt(rho_x ** 2 + rho_y ** 2)
    Gmax = np.max(G) + 1e-16
    alpha = 0.10
    S = (G > alpha * Gmax).astype(np.uint8)
    # :14-19 thin to curve
    C = skeletonize(S > 0)
    # :22 band
    band = binary_dilation(S, structure=np.ones((5, 5)))
    # :25-34 curve irregularity
    ys, xs = np.where(C)
    if len(xs) < 10:
        I_curve = 0.0
    else:
        A = np.vstack([ys, np.ones_like(ys)]).T
        a, b = np.linalg.lstsq(A, xs, rcond=None)[0]
        xs_fit = a * ys + b
        I_curve = np.std(xs - xs_fit)
    # :37-45 transverse energy
    nx = rho_x
    ny = rho_y
    mag = np.sqrt(nx * nx + ny * ny) + 1e-16
    nxn = nx / mag
    nyn = ny / mag
    tx = -nyn
    ty = nxn
    vperp = u * tx + v * ty
    E_perp = np.sum((vperp[band]) ** 2)
    # :48-55 overshoot
    smooth = median_filter(rho, size=7)
    delta = rho - smooth
    if np.any(band):
        O = np.max(np.abs(delta[band])) / (np.max(rho) - np.min(rho) + 1e-12)
    else:
        O = 0.0
    # :58-59 total variation
    dx_r = np.diff(rho, axis=0)
    dy_r = np.diff(rho, axis=1)
    TV = (np.sum(np.abs(dx_r)) + np.sum(np.abs(dy_r))) / (rho.size)
    # :62-64 scales
    I0 = max(1.0, np.sqrt(rho.shape[0]))
    E0 = max(1e-8, np.median(np.abs(v[band])) ** 2 * np.sum(band)) if np.any(band) else 1.0
    TV0 = max(1e-8, np.mean(np.abs(rho)))
    cI = I_curve / I0
    cE = E_perp / E0
    cO = O
    cTV = TV / TV0
    if weights is None:
        weights = {"I": 1.0, "E": 1.0, "O": 50.0, "TV": 1.0}
    J = weights["I"] * cI + weights["E"] * cE + weights["O"] * cO + weights["TV"] * cTV
    metrics = {"I_curve": I_curve, "E_perp": E_perp, "O": O, "TV": TV,
               "cI": cI, "cE": cE, "cO": cO, "cTV": cTV}
    return float(J), metrics


def objective_fields(rho, dx=1.0, dy=1.0, skeletonize=zhang_suen):
    """The integer and selection intermediates of the objective: shock mask S,
    skeleton C, band, and the 7 x 7 median of rho."""
    rho_x = (np.pad(rho, ((1, 1), (0, 0)), "edge")[2:, :] - np.pad(rho, ((1, 1), (0, 0)), "edge")[:-2, :]) / (2 * dx)
    rho_y = (np.pad(rho, ((0, 0), (1, 1)), "edge")[:, 2:] - np.pad(rho, ((0, 0), (1, 1)), "edge")[:, :-2]) / (2 * dy)
    G = np.sqrt(rho_x ** 2 + rho_y ** 2)
    S = G > 0.10 * (np.max(G) + 1e-16)
    return S, skeletonize(S), binary_dilation(S, structure=np.ones((5, 5))), median_filter(rho, size=7)
