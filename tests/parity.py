"""Parity checks shared by the GPU tests: every comparison prints its family, its largest scaled
error and the tolerance it is held to.

The default tolerance is 1e-5 relative.  Scales:
  * spectra: each multipole against its own magnitude, |x_i| floored at 1e-3 of that multipole's
    largest entry (bins where the multipole crosses zero);
  * Fisher matrices: sqrt(|F_ii F_jj|) (off-diagonal entries cross zero);
  * estimates that go through a Fisher inverse, x = F^-1 n: the norm-wise error is held to the
    perturbation bound kappa (e_n + e_F) / (1 - kappa e_F) (kappa = cond(F); e_n, e_F the measured
    relative errors of the numerator and of F), and each entry to 1e-5 of its multipole's own scale
    or, where larger, twice its componentwise bound |F^-1| (|dn| + |dF| |x|) from the same measured
    input errors.  Both bounds are computed in the test from the inputs it compares.
"""
import numpy as np

TOL = 1e-5


def _as_array(x):
    """float64, or complex128 for complex data (so that a comparison keeps the imaginary parts)."""
    x = np.asarray(x)
    return x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)


def scaled_error(got, ref, scale):
    got, ref = _as_array(got), _as_array(ref)
    return float(np.max(np.abs(got - ref) / np.maximum(np.asarray(scale, dtype=np.float64), 1e-300)))


def check(what, got, ref, scale=None, tol=TOL):
    """Assert max |got - ref| / scale < tol (scale = |ref| by default; complex data compared by the
    modulus of the difference) and print the error."""
    ref = _as_array(ref)
    s = np.abs(ref) if scale is None else scale
    err = scaled_error(got, ref, s)
    print("[parity] %s: max scaled error %.2e (tol %.1e)" % (what, err, tol), flush=True)
    assert err < tol, "%s: max scaled error %.3e >= %.1e" % (what, err, tol)
    return err


def own_scale(x, floor=1e-3):
    """|x_i| floored at floor * max |x| (one multipole's own magnitude)."""
    a = np.abs(np.asarray(x, dtype=np.float64))
    return np.maximum(a, floor * a.max())


def fisher_scale(F):
    d = np.abs(np.diag(np.asarray(F, dtype=np.float64)))
    return np.sqrt(np.outer(d, d))


def check_fisher(what, got, ref, tol=TOL):
    return check(what, got, ref, scale=fisher_scale(ref), tol=tol)


def solve_bound(F_ref, F_got, n_ref, n_got):
    """Bounds on the error of x = F^-1 n from the measured input errors dn = n_got - n_ref and
    dF = F_got - F_ref: the norm-wise kappa (e_n + e_F) / (1 - kappa e_F) (2-norms, kappa = cond(F)),
    and the componentwise first-order bound |F^-1| (|dn| + |dF| |x|) per entry."""
    F_ref, F_got = np.asarray(F_ref, dtype=np.float64), np.asarray(F_got, dtype=np.float64)
    n_ref, n_got = np.asarray(n_ref, dtype=np.float64), np.asarray(n_got, dtype=np.float64)
    kappa = np.linalg.cond(F_ref)
    eF = np.linalg.norm(F_got - F_ref, 2) / np.linalg.norm(F_ref, 2)
    en = np.linalg.norm(n_got - n_ref) / np.linalg.norm(n_ref)
    norm_bound = np.inf if kappa * eF >= 1 else kappa * (en + eF) / (1. - kappa * eF)
    Finv = np.linalg.inv(F_ref)
    x = Finv @ n_ref
    comp = np.abs(Finv) @ (np.abs(n_got - n_ref) + np.abs(F_got - F_ref) @ np.abs(x))
    return norm_bound, comp, kappa, en, eF


def check_solve(what, got, ref, keys, F_ref, F_got, n_ref, n_got):
    """Estimates x = F^-1 n (dicts of multipoles, `keys` in the order of n): the norm-wise error is
    held to the condition-number bound, and each entry to max(1e-5 x its multipole's own scale,
    2 x its componentwise bound), so that an entry may exceed 1e-5 of its own magnitude only where
    the measured numerator and Fisher errors, propagated through |F^-1|, already allow it."""
    norm_bound, comp, kappa, en, eF = solve_bound(F_ref, F_got, n_ref, n_got)
    xg = np.concatenate([np.asarray(got[k], dtype=np.float64) for k in keys])
    xr = np.concatenate([np.asarray(ref[k], dtype=np.float64) for k in keys])
    norm_err = np.linalg.norm(xg - xr) / np.linalg.norm(xr)
    print("[parity] %s: cond(F) %.1f, numerator error %.2e, Fisher error %.2e, norm-wise error %.2e "
          "(bound %.2e)" % (what, kappa, en, eF, norm_err, norm_bound), flush=True)
    assert norm_err <= 1.01 * norm_bound + 1e-15, "%s: norm-wise error %.3e above the bound %.3e" % (
        what, norm_err, norm_bound)
    scale = np.concatenate([own_scale(ref[k]) for k in keys])
    allowed = np.maximum(TOL * scale, 2. * comp + 1e-14 * np.abs(xr).max())
    err = np.abs(xg - xr)
    over = err > TOL * scale
    print("[parity] %s: max error %.2e of the own scale; %d of %d entries above 1e-5 of it, each within 2x "
          "its componentwise bound (max ratio %.2f)" % (what, float(np.max(err / scale)), int(over.sum()), len(err),
                                                        float(np.max(err / allowed))), flush=True)
    assert np.all(err <= allowed), "%s: entries beyond max(1e-5 scale, 2 x componentwise bound)" % what
    return norm_bound
