"""Comparison helpers shared by the GPU parity tests (no test cases here)."""

import numpy as np

# Integer and flag fields of a history row (jgs2-shader.ts:4421-4484):
# validity bits, accept/revert/failures, energy-pair validity, the two gate
# flags, converged/finite/row index, backtracks/skipped/evaluations.
HISTORY_INT_FIELDS = {(0, 3), (1, 0), (1, 1), (1, 2), (1, 3), (2, 3), (5, 2), (5, 3), (6, 0), (6, 1), (6, 2),
                      (7, 0), (7, 1), (7, 2), (7, 3)}
CONTROL_INT_FIELDS = {(0, 3), (1, 0), (1, 1), (1, 2), (1, 3), (2, 3), (3, 1), (3, 2), (3, 3)}
# Fields whose f32 value is a sum or difference of nearly cancelling terms
# carry an absolute rounding error at the scale of those terms, not of the
# result: the assembled energies (a frame's energies fall from tens to below
# one while the inertia term 1/2 m/h^2 |x - x~|^2 differences positions
# about 1e-3 |x| apart), the convergence gradient component norms, the
# summed-gradient norm and the denominator (a vertex's elastic force sums
# incident-tetrahedron forces that cancel near equilibrium), the relative
# residual, and the update norms (differences of two f32 positions). They
# are compared against both oracles relative to the scale of their terms:
# the frame's largest energy, the frame's largest residual denominator (the
# reference's own gradient scale), 1 for the relative residual, and the
# frame's largest update.
CANCELLING_TOL = 3e-4
ENERGY_FIELDS = {(2, 0), (2, 1), (2, 2)}
NORM_FIELDS = {(3, 0), (3, 1), (3, 2), (3, 3), (4, 0), (4, 1), (4, 2)}
UPDATE_FIELDS = {(5, 0), (5, 1)}
HISTORY_CANCELLING_FIELDS = ENERGY_FIELDS | NORM_FIELDS | UPDATE_FIELDS | {(4, 3)}
CONTROL_CANCELLING_FIELDS = ENERGY_FIELDS


F32_ULP = 2.0 ** -23


def cancelling_scales(history_rows, max_coordinate, scene_scale):
    """(scale, absolute floor) of each cancelling field from a frame's
    history rows (the FP64 oracle's). Updates are differences of f32
    positions, known to a few units in the last place of the largest
    coordinate; energies at rest are zero up to f64 rounding."""
    h = np.asarray(history_rows, np.float64).reshape(-1, 8, 4)
    ulp = 8.0 * F32_ULP * max_coordinate
    energy = np.abs(h[:, 2, :3]).max()
    denominator = np.abs(h[:, 4, 2]).max()
    scales = {f: (energy, 1e-12) for f in ENERGY_FIELDS}
    scales.update({f: (denominator, 1e-12) for f in NORM_FIELDS})
    scales[(5, 0)] = (np.abs(h[:, 5, 0]).max(), ulp)
    scales[(5, 1)] = (np.abs(h[:, 5, 1]).max(), ulp / max(scene_scale, 1e-12))
    scales[(4, 3)] = (1.0, 0.0)
    return scales


def compare_records(got, want, int_fields, tol, label, fp64=None, scales=None):
    """Integer fields equal; float fields within `tol` of the FP32 oracle
    (`want`) relative to the largest magnitude the field takes over the
    records (minimum determinants, shifts). Fields with an entry (scale,
    floor) in `scales` are held to max(CANCELLING_TOL * scale, floor) against
    the FP32 oracle and the FP64 oracle `fp64`. Returns the worst relative
    error of the `tol` fields."""
    scales = scales or {}
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    assert got.shape == want.shape, label
    rows, cols = got.shape[-2:]
    g = got.reshape(-1, rows, cols)
    refs = [want.reshape(-1, rows, cols)]
    if fp64 is not None:
        refs.append(np.asarray(fp64, np.float64).reshape(-1, rows, cols))
    worst = 0.0
    for r in range(rows):
        for c in range(cols):
            a, b = g[:, r, c], refs[0][:, r, c]
            if (r, c) in int_fields:
                bad = np.flatnonzero(a != b)
                assert bad.size == 0, f"{label}: field ({r},{c}) differs at records {bad[:8]}: {a[bad[:4]]} vs {b[bad[:4]]}"
                continue
            cancelling = (r, c) in scales
            for which, ref in enumerate(refs if cancelling else refs[:1]):
                e = ref[:, r, c]
                if cancelling:
                    scale, floor = scales[(r, c)]
                    allowed = max(CANCELLING_TOL * scale, floor)
                else:
                    allowed = tol * max(np.abs(e).max(), 1e-30)
                    worst = max(worst, np.abs(a - e).max() / max(np.abs(e).max(), 1e-30))
                err = np.abs(a - e).max()
                assert err <= allowed, (f"{label}: field ({r},{c}) error {err:.3e} > {allowed:.3e} against the "
                                        f"FP{32 if which == 0 else 64} oracle ({a[:3]} vs {e[:3]})")
    return worst


def rel_l2(a, b, ref=None):
    ref = b if ref is None else ref
    return float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)) /
                 max(np.linalg.norm(np.asarray(ref, np.float64)), 1e-30))


def xyz(p):
    return np.asarray(p, np.float64).reshape(-1, 4)[:, :3]


def local_system_errors(g_gpu, h_gpu, g_ref, h_ref, free):
    """Per-vertex errors of (g, H): H relative to the vertex's largest Hessian
    entry; g relative to the largest |g| over the free vertices (norm-wise),
    since a vertex near equilibrium has |g| far below its terms."""
    gs = np.linalg.norm(g_ref[free], axis=1).max()
    g_err = np.linalg.norm(g_gpu[free] - g_ref[free], axis=1) / gs
    h_err = np.abs(h_gpu[free] - h_ref[free]).max(axis=(1, 2)) / np.abs(h_ref[free]).max(axis=(1, 2))
    return float(g_err.max()), float(h_err.max())


def check_local_systems(g_gpu, h_gpu, systems, free, tol=1e-5, label=""):
    """The local-system bar. `systems` maps 32 and 64 to the FP32 and FP64
    oracles' (g, H). Each of g and H is within `tol` of both oracles where
    FP32 resolves it to that level (the FP32 oracle within `tol` of FP64);
    where FP32 itself loses digits (cancellation in g, the barrier Hessian
    near the activation distance), within three times the FP32 oracle's own
    distance from FP64. Returns the measured errors."""
    g32, h32 = systems[32]
    g64, h64 = systems[64]
    out = {}
    out["g_vs_32"], out["h_vs_32"] = local_system_errors(g_gpu, h_gpu, g32, h32, free)
    out["g_vs_64"], out["h_vs_64"] = local_system_errors(g_gpu, h_gpu, g64, h64, free)
    out["g_32_vs_64"], out["h_32_vs_64"] = local_system_errors(g32, h32, g64, h64, free)
    print(f"{label} g: vs FP32 {out['g_vs_32']:.2e}, vs FP64 {out['g_vs_64']:.2e} (FP32 oracle vs FP64 "
          f"{out['g_32_vs_64']:.2e}); H: vs FP32 {out['h_vs_32']:.2e}, vs FP64 {out['h_vs_64']:.2e} (FP32 oracle "
          f"vs FP64 {out['h_32_vs_64']:.2e})", flush=True)
    for q in ("g", "h"):
        if out[f"{q}_32_vs_64"] < tol:
            assert out[f"{q}_vs_32"] < tol and out[f"{q}_vs_64"] < tol, q
        else:
            assert out[f"{q}_vs_64"] <= 3.0 * out[f"{q}_32_vs_64"], q
    return out
