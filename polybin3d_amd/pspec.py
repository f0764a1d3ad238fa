"""Power-spectrum estimators, backend='hip' (PolyBin3D/pspec.py).

Same API as PolyBin3D.PSpec: Pk_ideal, Pk_numerator_ideal, compute_fisher_ideal, Pk_numerator,
compute_fisher_contribution, compute_fisher, compute_shot_contribution, Pk_unwindowed, get_ks.
Maps stay on the GPU; each step is a fused HIP kernel:
  * numerators: one R2C with the 1/P_fid weight in its epilogue, then one segmented bin reduction
    for all multipoles (global) or R2C-with-Y_lm prologue/epilogue harmonics sums (local);
  * MC Fisher: all Q maps of one multipole are produced by one batched filtered C2R (Theta_b L_l in
    the prologue, mask in the epilogue) followed by one batched R2C; the Fisher rows of all of them
    come from one bin-reduction launch.  Fisher rows accumulate in FP64.
Results are returned as numpy FP64 arrays like the reference's.
"""
import numpy as np
import scipy.interpolate
import torch

from . import _capi as C
from . import tune
from .base import harmonics_tags
from .constraints import Pointing, covariance_filter


def fp64_inverse(F, device):
    """inv(F) in FP64 on the GPU (LU), returned as a host array: the reference's np.linalg.inv of a given
    Fisher matrix (pspec.py:583, bspec.py:875), seconds on the host for 12 700 bins."""
    Fd = torch.as_tensor(np.ascontiguousarray(F, dtype=np.float64), device=device)
    return torch.linalg.inv(Fd).cpu().numpy()


def derivative_coefficient(edges):
    """Per-bin weights of the wide-angle k dP/dk term on the theory bin edges e_0..e_n (pspec.py:369-379):
    w_j = e_{j-1} / (e_j - e_{j-2}) [2 <= j]  -  e_{j+1} / (e_{j+2} - e_j) [j <= n-3]
          + e_0 / (e_1 - e_0) ([j = 1] - [j = 0])  -  e_n / (e_n - e_{n-1}) ([j = n-1] - [j = n-2])."""
    e = np.asarray(edges, dtype=np.float64)
    n = len(e) - 1
    j = np.arange(n)
    w = np.zeros(n)
    lo = j >= 2
    w[lo] += e[j[lo] - 1] / (e[j[lo]] - e[j[lo] - 2])
    hi = j <= n - 3
    w[hi] -= e[j[hi] + 1] / (e[j[hi] + 2] - e[j[hi]])
    first, last = e[0] / (e[1] - e[0]), e[n] / (e[n] - e[n - 1])
    w[0] -= first
    w[n - 1] -= last
    if n >= 2:
        w[1] += first
        w[n - 2] += last
    return w


class PSpec:
    """Power spectrum estimators (PolyBin3D/pspec.py:10-784)."""

    def __init__(self, base, k_bins, lmax=2, mask=None, applySinv=None, odd_l=False, add_GIC=False,
                 mask_IC=None, add_RIC=False, radial_bins_RIC=[], mask_shot=None, applySinv_transpose=None,
                 verb=False):
        # pspec.py:27-120
        self.base = base
        self.k_bins = np.asarray(k_bins, dtype=np.float64)
        self.lmax = lmax
        self.Nk = len(k_bins) - 1
        self.odd_l = odd_l
        self.Nl = lmax + 1 if odd_l else lmax // 2 + 1
        self.Nl_even = lmax // 2 + 1
        self.N_bins = self.Nk * self.Nl
        self.add_GIC, self.add_RIC = add_GIC, add_RIC
        assert np.max(self.k_bins) < np.min(base.kNy), "k_max must be less than k_Nyquist!"
        assert np.min(self.k_bins) >= np.max(base.kF), "k_min must be at least the k_fundamental!"
        assert np.max(self.k_bins) <= base.Pfid[0][-1], "Fiducial power spectrum should extend up to k_max!"
        assert np.min(self.k_bins) >= base.Pfid[0][0], "Fiducial power spectrum should extend down to k_min!"
        if odd_l:
            assert self.lmax > 0, "lmax must be greater than 0 to use odd harmonics!"
            assert base.sightline == 'local', "Odd spherical harmonics should only be used for local line-of-sight!"
        assert type(self.lmax) == int, "l-max must be an integer!"
        assert self.lmax <= 4, "Only l-max<=4 is currently supported!"
        if mask is None:
            self.const_mask, self.mask_mean, self.sq_mask_mean = True, 1., 1.
            assert not add_GIC, "Global integral constraint cannot be used without a mask!"
            assert not add_RIC, "Radial integral constraint cannot be used without a mask!"
        else:
            m = mask.detach().cpu().numpy() if torch.is_tensor(mask) else np.asarray(mask)
            m = np.ascontiguousarray(np.broadcast_to(m.astype(np.float64), tuple(int(v) for v in base.gridsize)))
            if np.std(m) < 1e-12:
                self.const_mask, self.mask_mean = True, float(np.mean(m))
            else:
                self.const_mask = False
            self.sq_mask_mean = float(np.mean(m ** 2))
            self.mask = base.as_real(m)
        if mask_shot is not None:
            ms = mask_shot.detach().cpu().numpy() if torch.is_tensor(mask_shot) else np.asarray(mask_shot)
            self.mask_shot = base.as_real(ms.astype(np.float64))
        # integral constraints (pspec.py:101-109): the pointing matrix of the varying mask
        if add_GIC:
            assert mask_IC is not None, "Need to supply mask_IC!"
        if add_RIC:
            assert not add_GIC, "Radial integral constraint imposes global integral constraint automatically!"
            assert mask_IC is not None, "Need to supply mask_IC!"
            assert len(radial_bins_RIC) > 0, "Radial bins need to be supplied for radial integral constraint!"
        if mask is not None and not self.const_mask:
            mic = None if mask_IC is None else (mask_IC.detach().cpu().numpy() if torch.is_tensor(mask_IC)
                                               else np.asarray(mask_IC, dtype=np.float64))
            self._P = Pointing(base, m, mic, add_GIC, add_RIC, radial_bins_RIC)
        # pspec.py:92-99
        self.user_Sinv = applySinv is not None
        if applySinv is None:
            self.applySinv = base.applySinv_trivial
        else:
            self.applySinv = base.wrap_user_filter(applySinv)
            if not callable(applySinv_transpose):
                applySinv_transpose = applySinv
        self.applySinv_transpose = (base.applySinv_trivial if applySinv_transpose is None
                                    else base.wrap_user_filter(applySinv_transpose))
        if base.sightline == 'local' and self.lmax > 0:
            self.Ylm_real = harmonics_tags(lmax, odd_l, 'real')
            self.Ylm_fourier = harmonics_tags(lmax, odd_l, 'fourier')
        else:
            self.Ylm_real = self.Ylm_fourier = None
        self.table = base.bin_table(self.k_bins)

    def _harm(self, real_map, ell, out=None):
        """sum_m FFT[m(x) Y_lm(x)] Y_lm(k) (apply_fourier_harmonics), times -i for odd l: the phase
        the reference gives odd-l legs (pspec.py:469-481, 542-543, 677-678); in a cross-integral
        Re[Q conj(-i leg)] it is the reference's Re[leg conj(i Q)] (pspec.py:410-411)."""
        b = self.base
        out = b.apply_fourier_harmonics(real_map, self.Ylm_real[ell], self.Ylm_fourier[ell], out=out)
        if ell % 2 == 1:
            out.mul_(-1j)
        return out

    def get_ks(self):
        """Bin centres (pspec.py:122-128)."""
        return np.asarray([0.5 * (self.k_bins[b] + self.k_bins[b + 1]) for b in range(self.Nk)])

    # ------------------------------------------------------------------ numerators
    def _filtered_data(self, data, filtering):
        """S^-1 d in Fourier space (pspec.py:617-640)."""
        b = self.base
        if filtering == 'ideal':
            # invPk0 / sqrt(<mask^2>), and / W when a pixel window is deconvolved (pspec.py:627-638)
            w = b.invPk0_grid
            if b.pixel_window != 'none':
                w = self._ideal_pw_weight()
            elif b.invPk0_is_one:
                w = None
            op = C.kop(C.KOP_WEIGHT, w=w, scale=1. / np.sqrt(self.sq_mask_mean)) if w is not None else \
                C.kop(C.KOP_PLAIN, scale=1. / np.sqrt(self.sq_mask_mean))
            return b.r2c(b.as_real(data), kop=op)
        if b.pixel_window != 'none':
            df = b.r2c(b.as_real(data), kop=C.kop(C.KOP_WEIGHT, w=self._inv_pw()))
            return b.as_fourier(self.applySinv(df, input_type='fourier', output_type='fourier'))
        return b.as_fourier(self.applySinv(b.as_real(data), input_type='real', output_type='fourier'))

    def _inv_pw(self):
        b = self.base
        key = 'inv_pw'
        if key not in b._bufs:
            pw = b.pixel_window_grid_host
            w = np.zeros(pw.shape)
            w[pw != 0] = 1. / pw[pw != 0]
            b._bufs[key] = torch.from_numpy(w.astype(np.float32)).to(b.device)
        return b._bufs[key]

    def _ideal_pw_weight(self):
        b = self.base
        key = 'ideal_pw'
        if key not in b._bufs:
            pw = b.pixel_window_grid_host
            w = np.zeros(pw.shape)
            w[pw != 0] = b.invPk0_grid_host[pw != 0] / pw[pw != 0]
            b._bufs[key] = torch.from_numpy(w.astype(np.float32)).to(b.device)
        return b._bufs[key]

    def _compute_pk_numerator(self, data, filtering='ideal'):
        """pspec.py:613-686 (numpy FP64 result)."""
        return self._pk_numerator_dev(data, filtering).cpu().numpy()

    def _pk_numerator_dev(self, data, filtering='ideal'):
        """pspec.py:613-686 as an FP64 device tensor.  Global: one bin pass for all multipoles.  Local:
        S_real, then for each l>0 the harmonics sum (2l+1 fused R2Cs) and one cross-integral pass."""
        assert filtering in ['ideal', 'Sinv'], "Unknown filtering option supplied!"
        b = self.base
        Nl, lf = (self.Nl, 2 - self.odd_l) if filtering == 'Sinv' else (self.Nl_even, 2)
        ells = [2 * li for li in range(Nl)]
        stack = torch.empty((Nl,) + b.half_shape, dtype=torch.complex64, device=b.device)
        S = self._filtered_data(data, filtering)
        stack[0].copy_(S)
        if b.sightline == 'global' or self.lmax == 0:
            out = b.bin_reduce(stack[0], self.k_bins, ells)[0]
        else:
            S_real = b.c2r(stack[0])
            for li in range(1, Nl):
                self._harm(S_real, lf * li, out=stack[li])
            out = b.bin_reduce(stack[0], self.k_bins, [0] * Nl, y=stack, y_ostride=b.H)[0]
        return out.reshape(-1) * (0.5 * b.volume / float(b.gridsize.prod()) ** 2)

    def _ideal_weight_sq(self):
        """a(k)^2 for the ideal filter S = a(k) d(k) (a = invPk0 [/W] / sqrt(<mask^2>)), FP32 device
        half grid, or None when a is the constant 1/sqrt(<mask^2>)."""
        b = self.base
        if not hasattr(self, '_w2'):
            if b.pixel_window != 'none':
                pw = b.pixel_window_grid_host
                w = np.zeros(pw.shape)
                w[pw != 0] = b.invPk0_grid_host[pw != 0] / pw[pw != 0]
            elif b.invPk0_is_one:
                w = None
            else:
                w = b.invPk0_grid_host
            self._w2 = None if w is None else \
                torch.from_numpy((w ** 2 / self.sq_mask_mean).astype(np.float32)).to(b.device)
        return self._w2

    def _pk_ideal_adjoint(self, data, grad_pk):
        """d(grad_pk . Pk_ideal)/d(data) for a real map `data` (global sightline): with v = F^-T grad_pk,
        the gradient is (V/N) C2R[ a^2 d(k) sum_l v[l, b(k)] L_l(mu) (x2 on the Nyquist plane) ], one
        R2C and one fused-prologue C2R (no counterpart in the reference)."""
        b = self.base
        if b.sightline != 'global' and self.lmax > 0:
            raise NotImplementedError("the Pk_ideal adjoint is implemented for the global sightline")
        Nl, Nk = self.Nl_even, self.Nk
        Finv = torch.from_numpy(np.ascontiguousarray(self.inv_fish_ideal)).to(b.device)
        v = Finv.t() @ grad_pk.to(torch.float64).reshape(-1)
        scale = b.volume / float(b.gridsize.prod())
        w2 = self._ideal_weight_sq()
        if w2 is None:
            scale /= self.sq_mask_mean
        coef = (v.reshape(Nl, Nk) * scale).to(torch.float32).contiguous()
        Z = b.r2c(b.as_real(data))
        return b.c2r(Z, kop=C.kop(C.KOP_BINPOLY, bin_table=b.bin_table(self.k_bins), w=w2, ft=coef, nk=Nl, terms=Nk,
                                  support=b.support(self.k_bins[-1])))

    def Pk_numerator_ideal(self, data):
        """pspec.py:689-700."""
        if not self.const_mask:
            assert tuple(np.shape(data)) == tuple(self.mask.shape), "Data must have same shape as the mask!"
        return self._compute_pk_numerator(data, filtering='ideal')

    def Pk_numerator(self, data):
        """pspec.py:131-139."""
        if not self.const_mask:
            assert tuple(np.shape(data)) == tuple(self.mask.shape), "Data must have same shape as the mask!"
        return self._compute_pk_numerator(data, filtering='Sinv')

    # ------------------------------------------------------------------ ideal Fisher
    def compute_fisher_ideal(self, discreteness_correction=True, Pk_fid='default'):
        """pspec.py:702-752: block-diagonal 0.5 sum degk invPk0^2 L_la L_lb, all (la, lb) in one pass."""
        b = self.base
        Nk, Nl = self.Nk, self.Nl_even
        if not (isinstance(Pk_fid, str) and Pk_fid == 'default'):
            if Pk_fid is None or (isinstance(Pk_fid, str) and Pk_fid == 'none'):
                k_tmp = np.arange(np.min(b.kF) / 10., np.max(b.kNy) * 2, np.min(b.kF) / 10.)
                Pfid = np.asarray([k_tmp, 1. + 0. * k_tmp])
            else:
                assert len(Pk_fid) == 2, "Pk should contain k and P_0(k) columns"
                Pfid = np.asarray(Pk_fid)
            P = scipy.interpolate.interp1d(Pfid[0], Pfid[1], bounds_error=False, fill_value=0.)(b.modk_half)
            inv = np.zeros(P.shape)
            good = (P != 0) & np.isfinite(P)
            inv[good] = 1. / P[good]
            inv_d = torch.from_numpy(inv.astype(np.float32)).to(b.device)
        else:
            inv_d = b.invPk0_grid
        fish = np.zeros((Nl * Nk, Nl * Nk))
        if discreteness_correction:
            pairs = [(la, lb) for la in range(Nl) for lb in range(Nl)]
            r = b.bin_reduce(inv_d, self.k_bins, [2 * la for la, _ in pairs], ells2=[2 * lb for _, lb in pairs],
                             real=True)[0].cpu().numpy()
            for o, (la, lb) in enumerate(pairs):
                fish[la * Nk:(la + 1) * Nk, lb * Nk:(lb + 1) * Nk] = np.diag(0.5 * r[o])
        else:
            d = 0.5 * b.bin_reduce(inv_d, self.k_bins, [0], real=True)[0, 0].cpu().numpy()
            for la in range(Nl):
                fish[la * Nk:(la + 1) * Nk, la * Nk:(la + 1) * Nk] = np.diag(d) / (4. * la + 1.)
        self.fish_ideal = fish
        self.inv_fish_ideal = np.linalg.inv(fish)
        return fish

    def Pk_ideal(self, data, fish_ideal=[], discreteness_correction=True):
        """pspec.py:754-784 (keeps the reference's caching: an existing inverse Fisher is reused)."""
        if len(fish_ideal) != 0:
            self.fish_ideal = np.asarray(fish_ideal)
            self.inv_fish_ideal = np.linalg.inv(self.fish_ideal)
        if not hasattr(self, 'inv_fish_ideal'):
            self.compute_fisher_ideal(discreteness_correction=discreteness_correction)
        out = np.matmul(self.inv_fish_ideal, self.Pk_numerator_ideal(data))
        res, idx = {}, 0
        for l in range(0, self.lmax + 1, 2):
            res['p%s' % l] = out[idx:idx + self.Nk]
            idx += self.Nk
        return res

    # ------------------------------------------------------------------ MC Fisher
    def _q_maps(self, src, li, Q, kmax_out=None):
        """Q_{l,b} = (S^-1 P)^dag Theta_b L_l S^-1 P a for all bins b of multipole index li, written into
        Q[Nk, H] (pspec.py:444-488) with trivial S^-1, as one fused pointing chain:
          global / l=0 : Q = R2C(P * C2R(Theta_b L_l SinvP))
          local  l>0   : Q = 0.5 R2C(P * [C2R(Theta_b leg) + sum_m Y_m(x) C2R(Theta_b Y_m(k) SinvP)])
        (the reference applies (S^-1 P)^dag to both local half-terms separately; the operator is linear,
        so one R2C of the summed real map is the same).  Only modes with |k| below the last k-bin edge
        (or kmax_out, if larger) are written."""
        b = self.base
        Nk = self.Nk
        bins = self._bins_dev
        SinvP, leg, SinvP_mi = src
        ell = (2 - self.odd_l) * li
        sup = b.support(self.k_bins[-1])
        sup_out = b.support(max(self.k_bins[-1], kmax_out or 0.))
        common = dict(bin_table=self.table, bins=bins, x_batch_stride=0, support=sup)
        if b.sightline == 'global' or li == 0:
            subs = [(C.kop(C.KOP_FILTER, ell=ell, **common), SinvP, 0, 0, 1.0)]
        else:
            subs = [(C.kop(C.KOP_FILTER, ell=0, **common), leg[li], 0, 0, 0.5)]
            src_y = SinvP_mi if ell % 2 == 1 else SinvP  # -i on the odd-l Y_lm-weighted maps
            subs += [(C.kop(C.KOP_YLM, ell=ell, m=m, **common), src_y, ell, m, 0.5) for m in range(2 * ell + 1)]
        mask, scale, ic, finish = self._chain_pointing(Nk)
        b.pointing(subs, Nk, Q, mask=mask, scale=scale, kop_out=C.kop(C.KOP_PLAIN, support=sup_out), ic=ic)
        finish(Q)

    def _chain_pointing(self, batch):
        """Masking of the fused Q-map chain: P^T of the varying mask (with its integral constraint;
        pspec.py:291-304), or the constant mask's mean.  Returns (mask, scale, ic, finish)."""
        if self.const_mask:
            return None, self.mask_mean, None, (lambda out: out)
        full = lambda wz: self.base.r2c(wz, batch=wz.shape[0])  # noqa: E731
        mask, ic, finish = self._P.chain(True, batch, full, 'full')
        return mask, 1.0, ic, finish

    def compute_fisher_contribution(self, seed, verb=False):
        """One MC realization of the Fisher matrix (pspec.py:141-147, 215-493; compute_cov=False).
        Returns an FP64 numpy array [N_bins, N_bins]; row = Q-map (l, b), as in the reference."""
        return self._fisher_contribution_dev(seed).cpu().numpy()

    def _fisher_contribution_dev(self, seed, a_map_fourier=None, theory=None):
        """The MC Fisher contribution of one seed as an FP64 device tensor; with `theory` (from
        _theory_setup) also the theory binning matrix from the same Q maps."""
        b = self.base
        if self.user_Sinv:
            out = self._fisher_contribution_generic(seed, theory=theory)
            return tuple(torch.from_numpy(x) for x in out) if theory is not None else torch.from_numpy(out)
        Nk, Nl = self.Nk, self.Nl
        if not hasattr(self, '_bins_dev'):
            self._bins_dev = torch.arange(Nk, dtype=torch.int32, device=b.device)
        # pspec.py:266-268
        a = b.grf_half(seed=seed + int(1e7), output_type='fourier') if a_map_fourier is None else b.as_fourier(a_map_fourier)
        # A^-1 a (pspec.py:328)
        w = b.ainv_weight()
        Ainv = b.kmap(a, C.kop(C.KOP_WEIGHT, w=w)) if w is not None else a
        # S^-1 P a (pspec.py:329-332), trivial S^-1
        if self.const_mask:
            SinvP = b.kmap(a, C.kop(C.KOP_PLAIN, scale=self.mask_mean))
        else:
            SinvP = b.r2c(self._P.apply(b.c2r(a)))
        local = b.sightline == 'local' and self.lmax > 0
        lefts = torch.empty((Nl,) + b.half_shape, dtype=torch.complex64, device=b.device)
        lefts[0].copy_(Ainv)
        leg = {}
        Ainv_real = None
        if local:
            SinvP_real = b.c2r(SinvP)
            Ainv_real = b.c2r(Ainv)
            lf = 2 - self.odd_l
            for li in range(1, Nl):
                self._harm(Ainv_real, lf * li, out=lefts[li])
                leg[li] = self._harm(SinvP_real, lf * li)
        else:
            for li in range(1, Nl):
                lefts[li].copy_(Ainv)
        SinvP_mi = SinvP * (-1j) if self.odd_l else None
        ells = [0] * Nl if local else [2 * li for li in range(Nl)]
        Q = torch.empty((Nk,) + b.half_shape, dtype=torch.complex64, device=b.device)
        F = torch.zeros((self.N_bins, self.N_bins), dtype=torch.float64, device=b.device)
        scaling = 0.5 * b.volume / b.gridsize.prod() ** 2
        if theory is not None:
            th_lefts = self._theory_lefts(theory, Ainv, Ainv_real)
            Bm = torch.zeros((self.N_bins, theory['N_bins']), dtype=torch.float64, device=b.device)
        kmax_out = theory['edges'][-1] if theory is not None else None
        for li in range(Nl):
            self._q_maps((SinvP, leg, SinvP_mi), li, Q, kmax_out=kmax_out)
            r = b.bin_reduce(Q, self.k_bins, ells, y=lefts, y_qstride=0, y_ostride=b.H, nq=Nk)  # [Nk, Nl, Nk]
            F[li * Nk:(li + 1) * Nk, :] = r.reshape(Nk, Nl * Nk) * scaling
            if theory is not None:
                self._theory_add(theory, th_lefts, Q, Bm[li * Nk:(li + 1) * Nk], scaling)
        return (F, Bm) if theory is not None else F

    def _fisher_contributions_dev(self, seeds):
        """Sum over `seeds` of the MC Fisher contributions (pspec.py:215-493 for each seed), FP64 on the
        device.  The seeds go through every pass together: their GRFs, filters and Q maps are batch
        entries of the same launches (a sub-map input is shared by the Nk bin entries of its seed), so
        small grids fill the GPU.  Chunks of seeds keep the fused chain's work maps within 16 GiB."""
        b = self.base
        seeds = [int(s) for s in seeds]
        if self.user_Sinv or len(seeds) == 1:
            F = None
            for s in seeds:
                f = self._fisher_contribution_dev(s).to(b.device)
                F = f if F is None else F + f
            return F
        Nk, Nl, H, N = self.Nk, self.Nl, b.H, b.N
        local = b.sightline == 'local' and self.lmax > 0
        gmax = 2 * (2 * (Nl - 1)) + 2 if local else 1
        pz = min(b.half_shape[2], b.support(self.k_bins[-1])[2] + 1)  # z planes of a compact work map
        per_seed = (gmax + 1) * Nk * int(b.gridsize[0]) * int(b.gridsize[1]) * pz * 8
        chunk = max(1, min(len(seeds), (16 << 30) // per_seed))
        F = torch.zeros((self.N_bins, self.N_bins), dtype=torch.float64, device=b.device)
        for c0 in range(0, len(seeds), chunk):
            F += self._fisher_batch(seeds[c0:c0 + chunk])
        return F

    def _fisher_batch(self, seeds):
        b = self.base
        R, Nk, Nl, H, N = len(seeds), self.Nk, self.Nl, b.H, b.N
        if not hasattr(self, '_bins_rep'):
            self._bins_rep = {}
        if R not in self._bins_rep:
            self._bins_rep[R] = torch.arange(Nk, dtype=torch.int32, device=b.device).repeat(R)
        bins = self._bins_rep[R]
        a = torch.stack([b.grf_half(seed=s + int(1e7), output_type='fourier') for s in seeds])
        w = b.ainv_weight()
        Ainv = b.kmap(a, C.kop(C.KOP_WEIGHT, w=w, x_batch_stride=H), batch=R) if w is not None else a
        if self.const_mask:
            SinvP = b.kmap(a, C.kop(C.KOP_PLAIN, scale=self.mask_mean, x_batch_stride=H), batch=R)
        else:
            SinvP = b.r2c(self._P.apply(b.c2r(a, batch=R)), batch=R)
        local = b.sightline == 'local' and self.lmax > 0
        lefts = torch.empty((R, Nl) + b.half_shape, dtype=torch.complex64, device=b.device)
        lefts[:, 0].copy_(Ainv)
        leg = {}
        if local:
            SinvP_real = b.c2r(SinvP, batch=R)
            Ainv_real = b.c2r(Ainv, batch=R)
            lf = 2 - self.odd_l
            for li in range(1, Nl):
                leg[li] = torch.empty((R,) + b.half_shape, dtype=torch.complex64, device=b.device)
                for r in range(R):
                    self._harm(Ainv_real[r], lf * li, out=lefts[r, li])
                    self._harm(SinvP_real[r], lf * li, out=leg[li][r])
            del SinvP_real, Ainv_real
        else:
            for li in range(1, Nl):
                lefts[:, li].copy_(Ainv)
        ells = [0] * Nl if local else [2 * li for li in range(Nl)]
        Q = torch.empty((R * Nk,) + b.half_shape, dtype=torch.complex64, device=b.device)
        F = torch.zeros((self.N_bins, self.N_bins), dtype=torch.float64, device=b.device)
        scaling = 0.5 * b.volume / b.gridsize.prod() ** 2
        sup = b.support(self.k_bins[-1])
        common = dict(bin_table=self.table, bins=bins, x_batch_stride=H, x_batch_group=Nk, support=sup)
        SinvP_mi = SinvP * (-1j) if self.odd_l else None
        for li in range(Nl):
            ell = (2 - self.odd_l) * li
            # the Q maps of _q_maps for every seed: batch entry r*Nk + bin reads seed r's maps
            if not local or li == 0:
                subs = [(C.kop(C.KOP_FILTER, ell=ell, **common), SinvP, 0, 0, 1.0)]
            else:
                src_y = SinvP_mi if ell % 2 == 1 else SinvP
                subs = [(C.kop(C.KOP_FILTER, ell=0, **common), leg[li], 0, 0, 0.5)]
                subs += [(C.kop(C.KOP_YLM, ell=ell, m=m, **common), src_y, ell, m, 0.5) for m in range(2 * ell + 1)]
            mask, scale, ic, finish = self._chain_pointing(R * Nk)
            b.pointing(subs, R * Nk, Q, mask=mask, scale=scale, kop_out=C.kop(C.KOP_PLAIN, support=sup), ic=ic)
            finish(Q)
            for r in range(R):
                red = b.bin_reduce(Q[r * Nk:(r + 1) * Nk], self.k_bins, ells, y=lefts[r], y_qstride=0,
                                   y_ostride=H, nq=Nk)
                F[li * Nk:(li + 1) * Nk, :] += red.reshape(Nk, Nl * Nk) * scaling
        return F

    def _fisher_contribution_generic(self, seed, cov=None, theory=None):
        """Reference call sequence (pspec.py:215-493) with maps on the GPU, one Q map at a time: for a
        user-supplied S^-1, for the covariance (cov: the gridded spectrum of base.pk_grid; the filter
        S^-1 [P xi P^T + N] S^-dag, pspec.py:278-325) and, with `theory`, the theory binning matrix.
        Returns FP64 numpy arrays: F, or (F, binning matrix)."""
        b = self.base
        Nk, Nl = self.Nk, self.Nl
        F = np.zeros((self.N_bins, self.N_bins))
        Bm = np.zeros((self.N_bins, theory['N_bins'])) if theory is not None else None
        a = b.grf_half(seed=seed + int(1e7), output_type='fourier')
        if not self.const_mask:
            a_real = b.c2r(a)
        def cov_filter(m, it):
            # S^-1 [P xi P^T + N] S^-dag (pspec.py:278-289, 312-325)
            return covariance_filter(self, m, it, cov, getattr(self, 'mask_shot', None))

        def apply_filter(m):
            if cov is not None:
                return cov_filter(m, 'fourier' if self.const_mask else 'real')
            if self.const_mask:
                return self.mask_mean * b.as_fourier(self.applySinv(m, input_type='fourier', output_type='fourier'))
            return b.as_fourier(self.applySinv(self._P.apply(m), input_type='real', output_type='fourier'))

        def apply_filter_dagger(m, real=False):
            it = 'real' if real else 'fourier'
            if cov is not None:
                return cov_filter(m, it)
            if self.const_mask:
                return self.mask_mean * b.as_fourier(self.applySinv_transpose(m, input_type=it, output_type='fourier'))
            r = b.as_real(self.applySinv_transpose(m, input_type=it, output_type='real'))
            return b.r2c(self._P.apply(r, transpose=True))

        w = b.ainv_weight()
        Ainv = b.kmap(a, C.kop(C.KOP_WEIGHT, w=w)) if w is not None else a
        SinvP = apply_filter(a if self.const_mask else a_real)
        local = b.sightline == 'local' and self.lmax > 0
        lefts = torch.empty((Nl,) + b.half_shape, dtype=torch.complex64, device=b.device)
        lefts[0].copy_(Ainv)
        Ainv_real = None
        if local:
            SinvP_real = b.c2r(SinvP)
            Ainv_real = b.c2r(Ainv)
            for li in range(1, Nl):
                self._harm(Ainv_real, (2 - self.odd_l) * li, out=lefts[li])
        else:
            for li in range(1, Nl):
                lefts[li].copy_(Ainv)
        if theory is not None:
            th_lefts = self._theory_lefts(theory, Ainv, Ainv_real)
        ells = [0] * Nl if local else [2 * li for li in range(Nl)]
        scaling = 0.5 * b.volume / b.gridsize.prod() ** 2

        def add_row(row, Qm):
            r = b.bin_reduce(Qm, self.k_bins, ells, y=lefts, y_qstride=0, y_ostride=b.H)[0]
            F[row] = r.reshape(-1).cpu().numpy() * scaling
            if theory is not None:
                acc = torch.zeros((1, theory['N_bins']), dtype=torch.float64, device=b.device)
                self._theory_add(theory, th_lefts, Qm, acc, scaling)
                Bm[row] = acc[0].cpu().numpy()

        for li in range(Nl):
            ell = (2 - self.odd_l) * li
            if not local or li == 0:
                for ki in range(Nk):
                    km = b.map_utils.fourier_filter(SinvP, ell, self.k_bins[ki], self.k_bins[ki + 1])
                    add_row(li * Nk + ki, apply_filter_dagger(km))
            else:
                legm = self._harm(SinvP_real, ell)
                src_y = SinvP * (-1j) if ell % 2 == 1 else SinvP
                for ki in range(Nk):
                    t = b.bin_table([self.k_bins[ki], self.k_bins[ki + 1]])
                    km = b.kmap(legm, C.kop(C.KOP_FILTER, bin_table=t))
                    Qm = 0.5 * apply_filter_dagger(km)
                    real_map = torch.zeros(tuple(int(v) for v in b.gridsize), dtype=torch.float32, device=b.device)
                    for m in range(2 * ell + 1):
                        b.c2r(src_y, out=real_map, kop=C.kop(C.KOP_YLM, bin_table=t, ell=ell, m=m),
                              rop=C.rop(C.ROP_YLM, ell=ell, m=m, accumulate=1))
                    Qm = Qm + 0.5 * apply_filter_dagger(real_map, real=True)
                    add_row(li * Nk + ki, Qm)
        return (F, Bm) if theory is not None else F

    # ------------------------------------------------------------------ covariance and theory matrices
    def compute_covariance_contribution(self, seed, Pk_cov, verb=False):
        """One realization of the Gaussian covariance of the estimator for the theory spectrum
        Pk_cov = [k, P_0 (, P_2, P_4)] without shot noise (pspec.py:149-163): the MC-Fisher rows with
        S^-1 [P xi P^T + N] S^-dag as the filter, xi = base.apply_xi and N = mask_shot."""
        if not self.const_mask:
            assert hasattr(self, 'mask_shot'), "Must supply mask_shot to compute shot-noise contribution"
        assert len(Pk_cov) != 0, "Must provide fiducial power spectrum if computing covariance!"
        Pk_cov = np.asarray(Pk_cov, dtype=np.float64)
        assert Pk_cov.ndim == 2, "Pk should contain k and P_0, (and optionally P_2, P_4) columns"
        assert len(Pk_cov) in [2, 3, 4], "Pk should contain k and P_0, (and optionally P_2, P_4) columns"
        assert (len(Pk_cov) - 2) * 2 <= self.lmax, "Can't use higher multipoles than lmax"
        key = Pk_cov.tobytes()
        if getattr(self, '_cov_key', None) != key:
            self._cov_grid, self._cov_key = self.base.pk_grid(Pk_cov), key
        return self._fisher_contribution_generic(seed, cov=self._cov_grid)

    def compute_theory_contribution(self, seed, k_bins_theory, lmax_theory=None, include_wideangle=False,
                                    verb=False):
        """(Fisher contribution, rectangular theory binning matrix [N_bins, Nl_theory * Nk_theory]) from
        one realization (pspec.py:165-197), optionally with the wide-angle (odd-l) terms.  Both come from
        the same Q maps; the binning matrix is one more bin reduction per multipole against the theory
        legs (pspec.py:381-436)."""
        b = self.base
        k_bins_theory = np.asarray(k_bins_theory, dtype=np.float64)
        assert np.max(k_bins_theory) < np.min(b.kNy), "k_max must be less than k_Nyquist!"
        assert np.min(k_bins_theory) >= np.max(b.kF), "k_min must be at least the k_fundamental!"
        assert np.max(k_bins_theory) <= b.Pfid[0][-1], "Fiducial power spectrum should extend up to k_max!"
        assert np.min(k_bins_theory) >= b.Pfid[0][0], "Fiducial power spectrum should extend down to k_min!"
        if lmax_theory is None:
            lmax_theory = self.lmax
        # the data's multipoles are looked up among the theory legs (the reference fails otherwise)
        assert lmax_theory >= self.lmax, "Can't use larger lmax for data than for theory!"
        if include_wideangle:
            assert b.sightline != 'global', "Wide-angle effects require local lines-of-sight!"
            assert not self.odd_l, "Cannot include wide-angle effects for odd l!"
            assert lmax_theory > 1, "There are no wide-angle effects for the monopole!"
            assert lmax_theory <= 4, "Higher-order effects not yet implemented!"
        elif not self.odd_l:
            assert lmax_theory % 2 == 0, "Must use even lmax if not including wide-angle effects!"
        th = self._theory_setup(k_bins_theory, lmax_theory, include_wideangle)
        F, Bm = self._fisher_contribution_dev(seed, theory=th)
        return F.cpu().numpy(), Bm.cpu().numpy()

    def _theory_setup(self, edges, lmax_theory, include_wideangle):
        """The binning matrix as a list of terms (pspec.py:221-231, 354-436): each is a left map, its
        Legendre order in the bin reduction, the column block it adds to and an optional per-bin weight.
        Left maps: 'ainv' (A^-1 a), ('harm', l, phase) = phase sum_m FFT[Y_lm A^-1 a] Y_lm(k), and
        ('harm_r', l, phase) the same of A^-1 a / |r|.  The reference's phases: i Q on the odd-l
        cross-integrals (a -i on the leg) and i kfac-weights on the wide-angle legs (an i on the leg)."""
        b = self.base
        odd_th = True if include_wideangle else self.odd_l
        Nl_th = lmax_theory + 1 if (odd_th and not include_wideangle) else lmax_theory // 2 + 1
        Nk_th = len(edges) - 1
        terms = [('ainv', 0, 0, None)]
        if b.sightline == 'global':
            terms += [('ainv', 2 * li, li, None) for li in range(1, Nl_th)]
        elif not include_wideangle:
            lf = 2 - odd_th
            terms += [(('harm', lf * li, -1j if (lf * li) % 2 else 1.), 0, li, None) for li in range(1, Nl_th)]
        else:
            kfac = derivative_coefficient(edges)
            terms += [(('harm', l, 1.), 0, l // 2, None) for l in range(2, lmax_theory + 1, 2)]
            terms += [(('harm_r', 1, 1j), 0, 1, 3. / 5. * (3 + kfac))]
            if lmax_theory >= 3:
                terms += [(('harm_r', 3, 1j), 0, 1, 3. / 5. * (2 - kfac))]
            if lmax_theory >= 4:
                terms += [(('harm_r', 3, 1j), 0, 2, 10. / 9. * (5 + kfac))]
        dev = lambda c: None if c is None else torch.as_tensor(c, dtype=torch.float64, device=b.device)  # noqa: E731
        return dict(edges=edges, Nk=Nk_th, Nl=Nl_th, N_bins=Nl_th * Nk_th,
                    terms=[(src, lg, blk, dev(c)) for src, lg, blk, c in terms])

    def _theory_lefts(self, th, Ainv, Ainv_real):
        """The left maps of the theory terms, stacked [T, H] (equal sources share one transform)."""
        b = self.base
        T = len(th['terms'])
        L = torch.empty((T,) + b.half_shape, dtype=torch.complex64, device=b.device)
        done = {}
        for t, (src, _, _, _) in enumerate(th['terms']):
            if src in done:
                L[t].copy_(L[done[src]])
                continue
            done[src] = t
            if src == 'ainv':
                L[t].copy_(Ainv)
                continue
            kind, ell, phase = src
            if Ainv_real is None:
                Ainv_real = b.c2r(Ainv)
            r = Ainv_real
            if kind == 'harm_r':
                if not hasattr(self, '_inv_r'):
                    self._inv_r = b.as_real(1. / b.modr_grid)
                r = Ainv_real * self._inv_r
            b.apply_fourier_harmonics(r, ell, ell, out=L[t])
            if phase != 1.:
                L[t].mul_(phase)
        return L

    def _theory_add(self, th, lefts, Q, out, scaling):
        """out[q, blk * Nk_th + b] += scaling * w_b * sum_{k in b} degk Re[Q_q conj(left)] L_l(mu) over the
        terms, for the nq = out.shape[0] Q maps in Q."""
        b = self.base
        nq, Nk_th = out.shape[0], th['Nk']
        red = b.bin_reduce(Q, th['edges'], [lg for _, lg, _, _ in th['terms']], y=lefts, y_qstride=0,
                           y_ostride=b.H, nq=nq)  # [nq, T, Nk_th]
        for t, (_, _, blk, wt) in enumerate(th['terms']):
            r = red[:, t, :] * scaling
            out[:, blk * Nk_th:(blk + 1) * Nk_th] += r * wt if wt is not None else r

    def seeds_per_pass(self):
        """Seeds per batched MC Fisher pass from the autotuner (knob pk_batch: 1, 2, 4 or 8, timed per
        seed), or 1 when nothing is cached and tuning is off."""
        b = self.base
        if self.user_Sinv:
            return 1
        shape = (tuple(int(v) for v in b.gridsize), self.Nk, self.lmax, bool(self.odd_l), b.sightline,
                 bool(self.const_mask), getattr(self, '_P', None) is not None and self._P.kind is not None)

        def run(R):
            f = self._fisher_contributions_dev(list(range(10 ** 6, 10 ** 6 + R))) if R > 1 else \
                self._fisher_contribution_dev(10 ** 6)
            del f
        return tune.choose('pk_batch', shape, [1, 2, 4, 8], run, b.device, default=1, work=lambda R: R)

    def compute_fisher(self, N_it, verb=False, batch=1):
        """Mean of N_it realizations, accumulated in FP64 on the GPU, then inverted (pspec.py:553-569).
        batch > 1 (no counterpart in the reference) processes that many seeds per pass (same result up
        to FP64 summation order); batch='auto' takes the autotuner's choice (seeds_per_pass)."""
        fish = None
        if isinstance(batch, str) and batch == 'auto':
            batch = self.seeds_per_pass()
        batch = max(1, int(batch))
        for s0 in range(0, N_it, batch):
            if verb and (batch > 1 or s0 % 5 == 0):
                print("Computing Fisher contribution %d of %d" % (s0 + 1, N_it))
            seeds = list(range(s0, min(N_it, s0 + batch)))
            f = self._fisher_contributions_dev(seeds) if len(seeds) > 1 else self._fisher_contribution_dev(seeds[0])
            f = f.to(self.base.device)
            fish = f / N_it if fish is None else fish + f / N_it
        self.fish = fish.cpu().numpy()
        self.inv_fish = torch.linalg.inv(fish).cpu().numpy()
        return self.fish

    def compute_shot_contribution(self, seed):
        """Shot-noise numerator from one GRF drawn with 1/P_fid (pspec.py:495-551): the cross-integrals of
        S^-1 a (with its Y_lm legs for the local sightline) and S^-1 N A^-1 a, N = mask_shot (or the mask
        mean for a constant mask).  With the trivial S^-1 the filters are mode weights and mask products
        fused into the transforms; a user S^-1 is called on the maps as the reference does."""
        b = self.base
        if not self.const_mask:
            assert hasattr(self, 'mask_shot'), "Must supply mask_shot to compute shot-noise contribution"
        PkA = [b.Pfid[0], 1. / b.Pfid[1]]
        if getattr(self, '_invPkA', None) is None:
            # 1/P_A on the grid, interpolated once (the weight is cached by the array's identity)
            modk = b.modk_half
            PkA_grid = scipy.interpolate.interp1d(PkA[0], PkA[1], bounds_error=False,
                                                  fill_value=(PkA[1][0], PkA[1][-1]))(modk)
            invPkA = np.zeros_like(PkA_grid)
            invPkA[modk != 0] = 1. / PkA_grid[modk != 0]
            self._invPkA = invPkA
        a = b.grf_half(seed=seed + int(1e7), Pk_input=PkA, output_type='fourier')
        wA = b.ainv_weight(self._invPkA)
        AiA = b.kmap(a, C.kop(C.KOP_WEIGHT, w=wA))
        if self.user_Sinv:
            # S^-1 N A^-1 a and S^-1 a with the user's S^-1 (pspec.py:514-522)
            Sinv = lambda m, it: b.as_fourier(self.applySinv(m, input_type=it, output_type='fourier'))  # noqa: E731
            if self.const_mask:
                SNA = Sinv(AiA, 'fourier') * self.mask_mean
            else:
                SNA = Sinv(b.c2r(AiA) * self.mask_shot, 'real')
            Sa = Sinv(a, 'fourier')
        elif self.const_mask:
            SNA = b.kmap(AiA, C.kop(C.KOP_PLAIN, scale=self.mask_mean))
            Sa = a
        else:
            AiA_r = b.c2r(AiA)
            SNA = b.r2c(None, rop=C.rop(C.ROP_MUL, a=AiA_r, b=self.mask_shot))
            Sa = a
        Nl = self.Nl
        local = self.lmax > 0 and b.sightline == 'local'
        stack = torch.empty((Nl,) + b.half_shape, dtype=torch.complex64, device=b.device)
        stack[0].copy_(Sa)
        if local:
            Sa_real = b.c2r(Sa)
            for li in range(1, Nl):
                self._harm(Sa_real, (2 - self.odd_l) * li, out=stack[li])
            r = b.bin_reduce(SNA, self.k_bins, [0] * Nl, y=stack, y_ostride=b.H)[0]
        else:
            r = b.bin_reduce(SNA, self.k_bins, [2 * li for li in range(Nl)], y=Sa, y_ostride=0)[0]
        shot = 0.5 * r.reshape(-1).cpu().numpy()
        return shot * b.volume / b.gridsize.prod() ** 2

    def Pk_unwindowed(self, data, fish=[], shot_num=[], subtract_shotnoise=False):
        """pspec.py:571-611 (a given Fisher matrix is inverted in FP64 on the GPU)."""
        if len(fish) != 0:
            self.fish = np.asarray(fish)
            self.inv_fish = fp64_inverse(self.fish, self.base.device)
        if len(shot_num) != 0 or subtract_shotnoise:
            self.shot_num = np.asarray(shot_num)
            self.shot = np.matmul(self.inv_fish, self.shot_num)
        if not hasattr(self, 'inv_fish'):
            raise Exception("Need to compute Fisher matrix first!")
        if not hasattr(self, 'shot') and subtract_shotnoise:
            raise Exception("Need to compute shot-noise first!")
        out = np.matmul(self.inv_fish, self.Pk_numerator(data))
        if subtract_shotnoise:
            out = out - self.shot
        res, idx = {}, 0
        for l in range(0, self.lmax + 1, 2 - self.odd_l):
            res['p%s' % l] = out[idx:idx + self.Nk]
            idx += self.Nk
        return res


def compute_fisher_mixed(specs, N_it, batch=1, verb=False):
    """MC Fisher matrices of several PSpec objects on different grids (e.g. 128^3 and 256^3)
    computed together: every spec's realizations (batch seeds per pass, as compute_fisher) are
    issued on a HIP stream of its own, so the GPU overlaps the small grid's launches with the large
    grid's.  Returns the list of Fisher matrices, each equal to that spec's compute_fisher(N_it,
    batch=batch) up to FP64 summation order, and sets fish / inv_fish on each spec.  No counterpart
    in the reference, whose compute_fisher (pspec.py:553-569) runs one spec's seeds serially.  The
    specs must have distinct base objects (a base's work buffers serve one stream)."""
    specs = list(specs)
    if not specs:
        return []
    if len({id(ps.base) for ps in specs}) != len(specs):
        raise ValueError("compute_fisher_mixed: every spec needs its own PolyBin3D base")
    dev = specs[0].base.device
    if any(ps.base.device != dev for ps in specs):
        raise ValueError("compute_fisher_mixed: all specs must be on one device")
    main = torch.cuda.current_stream(dev)
    streams = [torch.cuda.Stream(device=dev) for _ in specs]  # default priority
    for st in streams:
        st.wait_stream(main)
    fish = [None] * len(specs)
    batch = max(1, int(batch))
    for s0 in range(0, N_it, batch):
        if verb:
            print("Computing Fisher contributions %d-%d of %d" % (s0 + 1, min(N_it, s0 + batch), N_it))
        seeds = list(range(s0, min(N_it, s0 + batch)))
        for i, (ps, st) in enumerate(zip(specs, streams)):
            with torch.cuda.stream(st):
                f = ps._fisher_contributions_dev(seeds) if len(seeds) > 1 else ps._fisher_contribution_dev(seeds[0])
                f = f.to(dev) / N_it
                fish[i] = f if fish[i] is None else fish[i] + f
    for st in streams:
        main.wait_stream(st)
    out = []
    for ps, f in zip(specs, fish):
        ps.fish = f.cpu().numpy()
        ps.inv_fish = torch.linalg.inv(f).cpu().numpy()
        out.append(ps.fish)
    return out
