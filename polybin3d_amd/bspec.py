"""Bispectrum estimators, backend='hip' (PolyBin3D/bspec.py).

Same API as PolyBin3D.BSpec: Bk_ideal, Bk_numerator_ideal, compute_fisher_ideal, Bk_numerator,
compute_fisher_contribution, compute_fisher, Bk_unwindowed, generate_sims, load_sims, get_ks.
  * g_{l,b}(x) maps: one batched C2R per multipole with Theta_b L_l in the prologue (all bins at once);
  * triplet sums sum_x g_a g_b g_c: pb3d_triplet_sum over the map stack;
  * MC Fisher: Q maps stored packed over the in-bin modes only, assembled inside the C2R prologue
    from packed FT[g_A g_B] maps, masked in the C2R epilogue / R2C prologue, and contracted with the
    S^-1-weighted maps as per-bin GEMMs; rows accumulate in FP64.
"""
import numpy as np
import scipy.interpolate
import torch

from . import _capi as C
from . import tune
from .base import harmonics_tags
from .constraints import Pointing, covariance_filter


def triangle_bins(k_bins, k_bins_squeeze, lmax, include_partial_triangles):
    """all_bins (PolyBin3D/bspec.py:67-75) with the triangle test of _check_bin (bspec.py:171-199)."""
    kb = np.asarray(k_bins, dtype=np.float64)
    ks = np.asarray(k_bins_squeeze, dtype=np.float64)
    Nk, Nks = len(kb) - 1, len(ks) - 1

    def check(b1, b2, b3):
        if include_partial_triangles:
            hi = ks[b1 + 1] + ks[b2 + 1]
            if b1 > b2:
                lo = ks[b1] - ks[b2 + 1]
            elif b1 == b2:
                lo = 0.
            else:
                lo = ks[b2] - ks[b1 + 1]
            return not (ks[b3] >= hi - 1e-10 or ks[b3 + 1] <= lo + 1e-10)
        k1 = 0.5 * (kb[b1] + kb[b1 + 1])
        k2 = 0.5 * (ks[b2] + ks[b2 + 1])
        k3 = 0.5 * (ks[b3] + ks[b3 + 1])
        return not (k3 < abs(k1 - k2) or k3 > k1 + k2)

    rows = [(b1, b2, b3, l) for l in range(0, lmax + 1, 2) for b1 in range(Nk) for b2 in range(b1, Nks)
            for b3 in range(b2, Nks) if check(b1, b2, b3)]
    return np.asarray(rows, dtype=np.int64).reshape(-1, 4)


class BSpec:
    """Bispectrum estimators (PolyBin3D/bspec.py:12-1119).

    Extra keywords: precision ('bf16x6', 'bf16x3' or 'fp32': the triplet and Fisher contractions) and
    chain_streams (HIP streams over which the MC Fisher's Q-map chains are dealt; 1 = all on the
    current stream).  Neither changes the result beyond FP32 rounding order."""

    def __init__(self, base, k_bins, lmax=0, mask=None, applySinv=None, k_bins_squeeze=None,
                 include_partial_triangles=False, add_GIC=False, mask_IC=None, add_RIC=False, radial_bins_RIC=[],
                 Pk_fid=None, mask_shot=None, applySinv_transpose=None, verb=False, precision='bf16x6',
                 chain_streams=4):
        b = self.base = base
        self.k_bins = np.asarray(k_bins, dtype=np.float64)
        self.lmax = lmax
        self.Nk = len(k_bins) - 1
        self.Nl = lmax // 2 + 1
        self.include_partial_triangles = include_partial_triangles
        self.Pk_fid = Pk_fid
        self.precision = precision
        self.chain_streams = max(1, int(chain_streams))
        assert self.lmax % 2 == 0, "l-max must be even!"
        assert np.max(self.k_bins) < np.min(b.kNy), "k_max must be less than k_Nyquist!"
        assert np.min(self.k_bins) >= np.max(b.kF), "k_min must be at least the k_fundamental!"
        assert np.max(self.k_bins) <= b.Pfid[0][-1], "Fiducial power spectrum should extend up to k_max!"
        assert np.min(self.k_bins) >= b.Pfid[0][0], "Fiducial power spectrum should extend down to k_min!"
        if k_bins_squeeze is not None:
            self.k_bins_squeeze = np.asarray(k_bins_squeeze, dtype=np.float64)
            self.Nk_squeeze = len(k_bins_squeeze) - 1
            assert (self.k_bins_squeeze[:self.Nk + 1] == self.k_bins).all(), "k_bins_squeeze must contain all non-squeezed bins!"
            assert np.max(self.k_bins_squeeze) < np.min(b.kNy), "k_max must be less than k_Nyquist!"
            assert np.min(self.k_bins_squeeze) >= np.max(b.kF), "k_min must be at least the k_fundamental!"
            assert np.max(self.k_bins_squeeze) <= b.Pfid[0][-1], "Fiducial power spectrum should extend up to k_max!"
        else:
            self.k_bins_squeeze = self.k_bins.copy()
            self.Nk_squeeze = self.Nk
        self.all_bins = triangle_bins(self.k_bins, self.k_bins_squeeze, lmax, include_partial_triangles)
        self.N_bins = len(self.all_bins)
        self.N3 = self.N_bins // self.Nl
        if verb:
            print("N_bins: %d" % self.N_bins)
        assert type(self.lmax) == int, "l-max must be an integer!"
        assert self.lmax <= 4, "Only l-max<=4 is currently supported!"
        if precision not in self._PREC:
            raise ValueError("precision must be one of %s" % sorted(self._PREC))
        if mask is None:
            self.const_mask, self.mask_mean, self.cube_mask_mean = True, 1., 1.
            assert not add_GIC, "Global integral constraint cannot be used without a mask!"
            assert not add_RIC, "Radial integral constraint cannot be used without a mask!"
        else:
            m = mask.detach().cpu().numpy() if torch.is_tensor(mask) else np.asarray(mask)
            m = np.broadcast_to(np.asarray(m, dtype=np.float64), tuple(int(v) for v in b.gridsize))
            if np.std(m) < 1e-12:
                self.const_mask, self.mask_mean = True, float(np.mean(m))
            else:
                self.const_mask = False
            self.cube_mask_mean = float(np.mean(m ** 3))
            self.mask = b.as_real(np.ascontiguousarray(m))
        # integral constraints (bspec.py:86-112): the pointing matrix of the varying mask
        if add_GIC:
            assert mask_IC is not None, "Need to supply mask_IC!"
        if add_RIC:
            assert not add_GIC, "Radial integral constraint imposes global integral constraint automatically!"
            assert mask_IC is not None, "Need to supply mask_IC!"
            assert len(radial_bins_RIC) > 0, "Radial bins need to be supplied for radial integral constraint!"
        if mask is not None and not self.const_mask:
            mic = None if mask_IC is None else (mask_IC.detach().cpu().numpy() if torch.is_tensor(mask_IC)
                                               else np.asarray(mask_IC, dtype=np.float64))
            self._P = Pointing(b, np.ascontiguousarray(m), mic, add_GIC, add_RIC, radial_bins_RIC)
        if mask_shot is not None:
            ms = mask_shot.detach().cpu().numpy() if torch.is_tensor(mask_shot) else np.asarray(mask_shot)
            self.mask_shot = b.as_real(ms.astype(np.float64))
        self.user_Sinv = applySinv is not None
        if applySinv is None:
            self.applySinv = b.applySinv_trivial
        else:
            self.applySinv = b.wrap_user_filter(applySinv)
            if not callable(applySinv_transpose):
                applySinv_transpose = applySinv
        self.applySinv_transpose = (b.applySinv_trivial if applySinv_transpose is None
                                    else b.wrap_user_filter(applySinv_transpose))
        # fiducial weighting (bspec.py:138-156); Pk_fid=None means unit weights with invPk0(0)=0
        if isinstance(Pk_fid, str) and Pk_fid == 'default':
            inv = b.invPk0_grid_host
        else:
            if Pk_fid is None:
                k_tmp = np.arange(np.min(b.kF) / 10., np.max(b.kNy) * 2, np.min(b.kF) / 10.)
                Pfid = np.asarray([k_tmp, 1. + 0. * k_tmp])
            else:
                assert len(Pk_fid) == 2, "Pk should contain k and P_0(k) columns"
                Pfid = np.asarray(Pk_fid)
            P = scipy.interpolate.interp1d(Pfid[0], Pfid[1], bounds_error=False, fill_value=0.)(b.modk_half)
            inv = np.zeros(P.shape)
            inv[P != 0] = 1. / P[P != 0]
        self.invPk0_host = inv
        self.invPk0 = torch.from_numpy(inv.astype(np.float32)).to(b.device)
        self.invPk0_c = torch.complex(self.invPk0, torch.zeros_like(self.invPk0))
        self.Ylm_real = harmonics_tags(lmax, False, 'real') if (b.sightline == 'local' and lmax > 0) else None
        self.Ylm_fourier = harmonics_tags(lmax, False, 'fourier') if lmax > 0 else None
        self.table = b.bin_table(self.k_bins_squeeze)
        self._bins_dev = torch.arange(self.Nk_squeeze, dtype=torch.int32, device=b.device)
        self._sup = b.support(self.k_bins_squeeze[-1])

    def get_ks(self):
        """Central (k1, k2, k3) of every triangle bin (bspec.py:201-220)."""
        kc = 0.5 * (self.k_bins[1:] + self.k_bins[:-1])
        ksc = 0.5 * (self.k_bins_squeeze[1:] + self.k_bins_squeeze[:-1])
        t = self.all_bins[:self.N3]
        return np.asarray([kc[t[:, 0]], ksc[t[:, 1]], ksc[t[:, 2]]])

    # ------------------------------------------------------------------ building blocks
    def _real_stack(self, nmaps):
        b = self.base
        return torch.empty((nmaps,) + tuple(int(v) for v in b.gridsize), dtype=torch.float32, device=b.device)

    def _filtered_maps(self, src, ell=0, ell2=0, out=None, w=None):
        """out[q] = C2R(Theta_q L_ell L_ell2 w src) for every squeezed bin q, one batched launch."""
        b = self.base
        Nq = self.Nk_squeeze
        if out is None:
            out = self._real_stack(Nq)
        b.c2r(src, out=out, batch=Nq, kop=C.kop(C.KOP_FILTER, bin_table=self.table, bins=self._bins_dev, ell=ell,
                                                  ell2=ell2, w=w, x_batch_stride=0, support=self._sup))
        return out

    def _g_maps(self, S, out=None):
        """g_{l,b}(x) for all l, b (bspec.py:931-956) -> stack [Nl*Nk_squeeze, n0, n1, n2]."""
        b = self.base
        Nq = self.Nk_squeeze
        if out is None:
            out = self._real_stack(self.Nl * Nq)
        self._filtered_maps(S, 0, out=out[0:Nq])
        if self.lmax > 0:
            if b.sightline == 'global':
                for l in range(2, self.lmax + 1, 2):
                    self._filtered_maps(S, l, out=out[(l // 2) * Nq:(l // 2 + 1) * Nq])
            else:
                S_real = b.c2r(S)
                for l in range(2, self.lmax + 1, 2):
                    leg = b.apply_fourier_harmonics(S_real, self.Ylm_real[l], self.Ylm_fourier[l])
                    self._filtered_maps(leg, 0, out=out[(l // 2) * Nq:(l // 2 + 1) * Nq])
        return out

    _PREC = {'fp32': 0, 'bf16x3': 1, 'bf16x6': 2}

    def _triplet(self, stack, i1, i2, i3):
        """FP64 numpy array of sum_x stack[i1] stack[i2] stack[i3] (assemble_b3 / sum_triplet,
        utils.pyx:357-366, 614-636).  precision 'bf16x3' / 'bf16x6': one split-BF16 MFMA GEMM
        T[a][p] = sum_x g_a g_b(p) g_c(p) over the distinct third maps a = i3 and pairs (i1, i2);
        'fp32': FP32 products with FP64 accumulation on the vector ALU."""
        prec = self._PREC[self.precision]
        if prec > 0 and stack.shape[0] <= 256:
            return self._triplet_mfma(stack, np.asarray(i1), np.asarray(i2), np.asarray(i3), prec)
        return self._triplet_vec(stack, i1, i2, i3)

    def _triplet_plan(self, nm, i1, i2, i3):
        """Launch list of the MFMA triplet sums: per block of at most 128 distinct third maps, the
        device index arrays (A maps, pair maps b <= c) and where each triplet reads its T entry."""
        b = self.base
        lo, hi = np.minimum(i1, i2), np.maximum(i1, i2)
        ua, ia = np.unique(i3, return_inverse=True)
        dev = lambda a, dt=np.int32: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(b.device)  # noqa: E731
        plan = []
        for c0 in range(0, len(ua), 128):
            sel = np.nonzero((ia >= c0) & (ia < c0 + 128))[0]
            up, ip = np.unique(lo[sel] * nm + hi[sel], return_inverse=True)
            amap = ua[c0:c0 + 128]
            plan.append((dev(amap), dev(up // nm), dev(up % nm), len(amap), len(up), dev(sel, np.int64),
                         dev(ia[sel] - c0, np.int64), dev(ip, np.int64)))
        return plan

    def _triplet_mfma(self, stack, i1, i2, i3, prec):
        b = self.base
        nm = stack.shape[0]
        nvox = int(np.prod(stack.shape[1:]))
        # launch lists are kept for read-only index arrays (the estimator's own triplet list), looked
        # up by identity; other index arrays get a fresh list
        plans = getattr(self, '_tplans', [])
        plan = next((e[4] for e in plans if e[0] is i1 and e[1] is i2 and e[2] is i3 and e[3] == nm), None)
        if plan is None:
            plan = self._triplet_plan(nm, i1, i2, i3)
            if not (i1.flags.writeable or i2.flags.writeable or i3.flags.writeable):
                self._tplans = (plans + [(i1, i2, i3, nm, plan)])[-4:]
        out = torch.empty(len(i1), dtype=torch.float64, device=b.device)
        for amap_d, pb_d, pc_d, na, npair, sel_d, rows_d, cols_d in plan:
            T = torch.empty((na, npair), dtype=torch.float64, device=b.device)
            ws = C.lib().pb3d_pair_gemm_work_size(nvox, na, npair)
            work = b.work('pair_gemm', ws)

            def launch(cfg, T=T, work=work, amap_d=amap_d, pb_d=pb_d, pc_d=pc_d, na=na, npair=npair):
                rc = C.lib().pb3d_pair_gemm_cfg(stack.data_ptr(), nm, nvox, amap_d.data_ptr(), na, pb_d.data_ptr(),
                                                pc_d.data_ptr(), npair, prec, cfg[0], cfg[1], T.data_ptr(),
                                                work.data_ptr(), b.stream())
                if rc == -1:
                    raise ValueError(C.lib().pb3d_last_error().decode())
                C.check(rc)
            cfg = tune.choose('pair_gemm', (nm, nvox, na, npair, prec), self._pair_gemm_shapes(nm, na), launch,
                              b.device, default=(0, 0))
            launch(cfg)
            out[sel_d] = T[rows_d, cols_d]
        return out.cpu().numpy()

    @staticmethod
    def _pair_gemm_shapes(nm, na):
        """Candidate (LDS stage, A rows per launch) shapes of pb3d_pair_gemm_cfg for a stack of nm maps
        and na A rows (0 = the built-in choice)."""
        long_stage = 128 if nm <= 32 else (64 if nm <= 128 else 32)
        stages = sorted({0, 32, long_stage})
        rows = [0] + [r for r in (32, 64, 96) if r < na]
        return [(st, r) for st in stages for r in rows]

    def _triplet_vec(self, stack, i1, i2, i3):
        b = self.base
        nm = stack.shape[0]
        nvox = int(np.prod(stack.shape[1:]))
        t = len(i1)
        idx = torch.from_numpy(np.stack([np.asarray(i1), np.asarray(i2), np.asarray(i3)]).astype(np.int32)).to(b.device)
        out = torch.empty(t, dtype=torch.float64, device=b.device)
        prec = 0
        ws = C.lib().pb3d_triplet_work_size(nvox, t, prec)
        work = b.work('triplet', ws)
        C.check(C.lib().pb3d_triplet_sum(stack.data_ptr(), nm, nvox, idx[0].data_ptr(), idx[1].data_ptr(),
                                         idx[2].data_ptr(), t, prec, out.data_ptr(), work.data_ptr(), b.stream()))
        return out.cpu().numpy()

    def _compute_symmetry_factor(self):
        """bspec.py:222-279: l=0 combinatorial; l>0 from u_b = IFFT(Theta_b invPk0) and
        sum_m IFFT(Theta_b Y_lm invPk0)^2 maps (batched C2R with a squaring epilogue)."""
        b = self.base
        Nq = self.Nk_squeeze
        t = self.all_bins[:self.N3]
        sym = [6. if (b1 == b2 == b3) else (2. if (b1 == b2 or b1 == b3 or b2 == b3) else 1.) for (b1, b2, b3, _) in t]
        if self.lmax > 0:
            stack = self._real_stack(2 * Nq + 1)
            self._filtered_maps(self.invPk0_c, 0, out=stack[0:Nq])
            stack[2 * Nq].fill_(1.)
            for l in range(2, self.lmax + 1, 2):
                b23 = stack[Nq:2 * Nq]
                for m in range(2 * l + 1):
                    b.c2r(self.invPk0_c, out=b23, batch=Nq,
                          kop=C.kop(C.KOP_YLM, bin_table=self.table, bins=self._bins_dev, ell=l, m=m, x_batch_stride=0),
                          rop=C.rop(C.ROP_SQUARE, accumulate=m > 0))
                one = 2 * Nq
                i1, i2, i3 = [], [], []
                for (b1, b2, b3, _) in t:
                    q = b1 if (b1 != b2 and b2 == b3) else b3
                    i1 += [q, q]
                    i2 += [Nq + b3, b3]
                    i3 += [one, b3]
                r = self._triplet(stack, i1, i2, i3).reshape(-1, 2)
                for i, (b1, b2, b3, _) in enumerate(t):
                    if b1 != b2 and b2 != b3:
                        sym.append(1.)
                    elif b1 == b2 and b2 != b3:
                        sym.append(2.)
                    elif b1 != b2 and b2 == b3:
                        sym.append(1. + r[i, 0] / r[i, 1])
                    else:
                        sym.append(2. * (1. + 2. * r[i, 0] / r[i, 1]))
            del stack
        self.sym_factor = np.asarray(sym)
        assert len(self.sym_factor) == self.N_bins

    # ------------------------------------------------------------------ numerators
    def _filtered_data(self, data, filtering):
        b = self.base
        if filtering == 'ideal':
            w = self.invPk0 if b.pixel_window == 'none' else self._filtered_weight_pw()
            return b.r2c(b.as_real(data), kop=C.kop(C.KOP_WEIGHT, w=w, scale=1. / self.cube_mask_mean ** (1. / 3.)))
        if b.pixel_window != 'none':
            pw = b.pixel_window_grid_host
            ww = np.zeros(pw.shape)
            ww[pw != 0] = 1. / pw[pw != 0]
            df = b.r2c(b.as_real(data), kop=C.kop(C.KOP_WEIGHT, w=torch.from_numpy(ww.astype(np.float32)).to(b.device)))
            return b.as_fourier(self.applySinv(df, input_type='fourier', output_type='fourier'))
        return b.as_fourier(self.applySinv(b.as_real(data), input_type='real', output_type='fourier'))

    def _b3_indices(self):
        """(b1, b2, stack row of b3 with its multipole) of every bin, computed once (read-only)."""
        if getattr(self, '_b3_idx', None) is None:
            Nq = self.Nk_squeeze
            ab = self.all_bins
            idx = (ab[:, 0].copy(), ab[:, 1].copy(), (ab[:, 3] // 2) * Nq + ab[:, 2])
            for v in idx:
                v.setflags(write=False)
            self._b3_idx = idx
        return self._b3_idx

    def _compute_bk_numerator(self, data, include_linear_term=True, filtering='ideal', verb=False):
        """bspec.py:895-983."""
        assert filtering in ['ideal', 'Sinv'], "Unknown filtering option supplied!"
        b = self.base
        if (not hasattr(self, 'preload')) and include_linear_term:
            raise Exception("Need to generate or specify bias simulations!")
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        S = self._filtered_data(data, filtering)
        g = self._g_maps(S)
        i1, i2, i3 = self._b3_indices()
        B3 = self._triplet(g, i1, i2, i3)
        B1 = np.zeros(self.N_bins)
        if include_linear_term:
            NG = g.shape[0]
            for ii in range(self.N_it):
                sim = self.g_bl_maps[ii] if self.preload else self.load_sim_data(ii)
                st = torch.cat([g, sim], 0)
                # g[0][b1] s[0][b2] s[l][b3] + s[0][b1] g[0][b2] s[l][b3] + s[0][b1] s[0][b2] g[l][b3]
                r = (self._triplet(st, i1, NG + i2, NG + i3) + self._triplet(st, NG + i1, i2, NG + i3)
                     + self._triplet(st, NG + i1, NG + i2, i3))
                B1 += -1. / self.N_it * r
                del st
        norm = b.volume / b.gridsize.prod() / self.sym_factor
        return (B3 + B1) * norm

    def Bk_numerator_ideal(self, data, verb=False):
        """bspec.py:986-997."""
        return self._compute_bk_numerator(data, include_linear_term=False, filtering='ideal', verb=verb)

    def Bk_numerator(self, data, include_linear_term=True, verb=False):
        """bspec.py:429-439."""
        return self._compute_bk_numerator(data, include_linear_term=include_linear_term, filtering='Sinv', verb=verb)

    # ------------------------------------------------------------------ ideal Fisher
    def compute_fisher_ideal(self, discreteness_correction=True, verb=False):
        """bspec.py:999-1083."""
        b = self.base
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        Nq, N3, Nl = self.Nk_squeeze, self.N3, self.Nl
        t = self.all_bins[:N3]
        if discreteness_correction:
            pairs = [(l, lp) for l in range(0, self.lmax + 1, 2) for lp in range(l, self.lmax + 1, 2)]
            stack = self._real_stack((Nl + len(pairs)) * Nq)
            for l in range(0, self.lmax + 1, 2):
                self._filtered_maps(self.invPk0_c, l, out=stack[(l // 2) * Nq:(l // 2 + 1) * Nq])
            fish = np.zeros((self.N_bins, self.N_bins))
            for p, (l, lp) in enumerate(pairs):
                o = (Nl + p) * Nq
                self._filtered_maps(self.invPk0_c, l, lp, out=stack[o:o + Nq])
                # case table of bspec.py:1038-1047: A = (u0[b1], u0[b2], u_llp[b3]),
                # B = (u0[b1], u_l[b2], u_lp[b3])
                i1 = list(t[:, 0]) * 2
                i2 = list(t[:, 1]) + list((l // 2) * Nq + t[:, 1])
                i3 = list(o + t[:, 2]) + list((lp // 2) * Nq + t[:, 2])
                r = self._triplet(stack, i1, i2, i3)
                A, B = r[:N3], r[N3:]
                for i, (b1, b2, b3, _) in enumerate(t):
                    if b1 != b2 and b2 != b3:
                        v = A[i]
                    elif b1 == b2 and b2 != b3:
                        v = 2. * A[i]
                    elif b1 != b2 and b2 == b3:
                        v = B[i] + A[i]
                    else:
                        v = 4. * B[i] + 2. * A[i]
                    fish[l // 2 * N3 + i, lp // 2 * N3 + i] = v
            del stack
            fish = fish + fish.T - np.diag(np.diag(fish))
            fish = fish / np.outer(self.sym_factor, self.sym_factor) * b.gridsize.prod() ** 2 / b.volume
        else:
            stack = self._filtered_maps(self.invPk0_c, 0)
            Ntri = self._triplet(stack, t[:, 0], t[:, 1], t[:, 2])
            del stack
            d = np.concatenate([Ntri / (2. * l + 1.) for l in range(0, self.lmax + 1, 2)])
            fish = np.diag(d / self.sym_factor) * b.gridsize.prod() ** 2 / b.volume
        self.fish_ideal = fish
        self.inv_fish_ideal = np.linalg.inv(fish)
        self.discreteness_correction = discreteness_correction
        return fish

    def Bk_ideal(self, data, fish_ideal=[], discreteness_correction=True, verb=False):
        """bspec.py:1085-1119."""
        if len(fish_ideal) != 0:
            self.fish_ideal = np.asarray(fish_ideal)
            self.inv_fish_ideal = np.linalg.inv(self.fish_ideal)
        if not hasattr(self, 'inv_fish_ideal'):
            self.compute_fisher_ideal(discreteness_correction=discreteness_correction, verb=verb)
        elif getattr(self, 'discreteness_correction', discreteness_correction) != discreteness_correction:
            self.compute_fisher_ideal(discreteness_correction=discreteness_correction, verb=verb)
        out = np.matmul(self.inv_fish_ideal, self.Bk_numerator_ideal(data, verb=verb))
        res, idx = {}, 0
        for l in range(0, self.lmax + 1, 2):
            res['b%s' % l] = out[idx:idx + self.N3]
            idx += self.N3
        return res

    def _bk_ideal_adjoint(self, data, grad_bk):
        """d(grad_bk . Bk_ideal)/d(data) for a real map `data` (global sightline; no counterpart in the
        reference).  With u = (F^-T grad_bk) * V/N/sym, the g-maps g_m = C2R(F_m a d) are recomputed,
        the adjoint maps Gamma_m = sum_i u_i d(g_i1 g_i2 g_i3)/d g_m come from one split-BF16 MFMA
        contraction, and the gradient is C2R( sum_m F_m a R2C(Gamma_m) ) with F_(l,b) = Theta_b L_l."""
        b = self.base
        if b.sightline != 'global' and self.lmax > 0:
            raise NotImplementedError("the Bk_ideal adjoint is implemented for the global sightline")
        Nq = self.Nk_squeeze
        Finv = torch.from_numpy(np.ascontiguousarray(self.inv_fish_ideal)).to(b.device)
        u = (Finv.t() @ grad_bk.to(torch.float64).reshape(-1)).cpu().numpy()
        u = u * b.volume / float(b.gridsize.prod()) / self.sym_factor
        i1, i2, i3 = [np.asarray(v, dtype=np.int64) for v in self._b3_indices()]
        Nm = self.Nl * Nq
        # U[m][pair] with Gamma_m = sum_pairs U[m][pair] g_b g_c
        rows = np.concatenate([i1, i2, i3])
        pa = np.concatenate([i2, i1, i1])
        pb_ = np.concatenate([i3, i3, i2])
        keys = np.minimum(pa, pb_) * Nm + np.maximum(pa, pb_)
        up, ip = np.unique(keys, return_inverse=True)
        U = np.zeros((Nm, len(up)))
        np.add.at(U, (rows, ip), np.concatenate([u, u, u]))
        S = self._filtered_data(data, 'ideal')
        g = self._g_maps(S)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(b.device)
        pb_d, pc_d = dev(up // Nm), dev(up % Nm)
        U_d = torch.from_numpy(U.astype(np.float32)).to(b.device)
        if Nm > 128:
            raise NotImplementedError("the Bk_ideal adjoint supports at most 128 g-maps (Nl * Nk_squeeze)")
        gam = torch.empty_like(g)
        prec = 1 if self.precision == 'bf16x3' else 2
        ws = C.lib().pb3d_pair_adjoint_work_size(Nm, len(up), prec)
        work = b.work('pair_adjoint', ws)
        C.check(C.lib().pb3d_pair_adjoint(g.data_ptr(), Nm, int(b.N), pb_d.data_ptr(), pc_d.data_ptr(), len(up),
                                          U_d.data_ptr(), Nm, prec, gam.data_ptr(), work.data_ptr(), b.stream()))
        del g
        w = self.invPk0
        if b.pixel_window != 'none':
            w = self._filtered_weight_pw()
        scale = 1. / self.cube_mask_mean ** (1. / 3.)
        Z = torch.zeros(b.half_shape, dtype=torch.complex64, device=b.device)
        for li in range(self.Nl):
            for q in range(Nq):
                b.r2c(gam[li * Nq + q], out=Z,
                      kop=C.kop(C.KOP_FILTER, bin_table=self.table, bins=self._bins_dev[q:q + 1], ell=2 * li, w=w,
                                scale=scale, accumulate=1))
        return b.c2r(Z)

    def _filtered_weight_pw(self):
        """invPk0 / W as a device FP32 half grid (the ideal filter with a pixel window), made once."""
        if getattr(self, '_w_pw', None) is None:
            b = self.base
            pw = b.pixel_window_grid_host
            ww = np.zeros(pw.shape)
            ww[pw != 0] = self.invPk0_host[pw != 0] / pw[pw != 0]
            self._w_pw = torch.from_numpy(ww.astype(np.float32)).to(b.device)
        return self._w_pw

    # ------------------------------------------------------------------ linear-term simulations
    def apply_pointing(self, map_real, transpose=False):
        """The pointing matrix of the varying mask, with its integral constraint (bspec.py:412-426)."""
        return self._P.apply(map_real, transpose)

    def _process_sim(self, sim, input_type='real'):
        """g_{b,l} maps of one simulation (bspec.py:281-326)."""
        b = self.base
        if b.pixel_window != 'none':
            sf = b.r2c(b.as_real(sim)) if input_type == 'real' else b.as_fourier(sim)
            pw = b.pixel_window_grid
            inv = torch.where(pw != 0, 1. / pw, torch.zeros_like(pw))
            sf = b.kmap(sf, C.kop(C.KOP_WEIGHT, w=inv.contiguous()))
            S = b.as_fourier(self.applySinv(sf, input_type='fourier', output_type='fourier'))
        else:
            S = b.as_fourier(self.applySinv(sim, input_type=input_type, output_type='fourier'))
        return self._g_maps(S)

    def load_sims(self, load_sim, N_sims, verb=False, input_type='real', preload=True):
        """bspec.py:328-360."""
        self.N_it = N_sims
        if preload:
            self.preload = True
            self.g_bl_maps = [self._process_sim(load_sim(ii), input_type=input_type) for ii in range(N_sims)]
        else:
            self.preload = False
            self.load_sim_data = lambda ii: self._process_sim(load_sim(ii), input_type=input_type)

    def generate_sims(self, N_sim, Pk_input=[], verb=False, preload=True):
        """bspec.py:362-410: GRFs with seeds 1e5+ii, pointing matrix and pixel window applied."""
        b = self.base
        self.N_it = N_sim

        def load_sim_data(ii):
            if self.const_mask:
                sf = b.grf_half(int(1e5) + ii, Pk_input=Pk_input, output_type='fourier')
                sf = b.kmap(sf, C.kop(C.KOP_PLAIN, scale=self.mask_mean))
                return self._process_sim(sf, input_type='fourier')
            sr = b.generate_data(int(1e5) + ii, Pk_input=Pk_input, output_type='real', include_pixel_window=False)
            if b.pixel_window != 'none':
                sf = b.r2c(self.apply_pointing(sr), kop=C.kop(C.KOP_WEIGHT, w=b.pixel_window_grid))
                return self._process_sim(sf, input_type='fourier')
            return self._process_sim(self.apply_pointing(sr), input_type='real')

        if preload:
            self.preload = True
            self.g_bl_maps = [load_sim_data(ii) for ii in range(N_sim)]
        else:
            self.preload = False
            self.load_sim_data = load_sim_data

    # ------------------------------------------------------------------ MC Fisher
    def _packed_layout(self):
        """In-bin modes of the squeezed bins, bin-major (segments of each bin contiguous):
        modes [M] (linear half-grid index), seg [Nq+1], pack_index [H] (-1 outside), degk [M]."""
        if getattr(self, '_pk', None) is None:
            b = self.base
            idx = b.bin_index_host(self.k_bins_squeeze).reshape(-1)
            lin = np.nonzero(idx >= 0)[0]
            order = np.argsort(idx[lin], kind='stable')
            modes = lin[order]
            seg = np.searchsorted(idx[modes], np.arange(self.Nk_squeeze + 1))
            pack_index = np.full(b.H, -1, dtype=np.int32)
            pack_index[modes] = np.arange(len(modes), dtype=np.int32)
            degk = np.where(modes % b.half_shape[2] > 0, 2., 1.)
            dev = b.device
            self._pk = dict(M=len(modes), seg=seg,
                            modes=torch.from_numpy(modes.astype(np.int32)).to(dev),
                            pack_index=torch.from_numpy(pack_index).to(dev),
                            degk=torch.from_numpy(degk.astype(np.float32)).to(dev))
        return self._pk

    def _pair_rows(self, pairs):
        """Row of each (A, B) pair in an FT table and the device [Nq*Nq] lookup (-1 = absent)."""
        Nq = self.Nk_squeeze
        rows = {p: r for r, p in enumerate(pairs)}
        tab = np.full(Nq * Nq, -1, dtype=np.int32)
        for (A, B), r in rows.items():
            tab[A * Nq + B] = r
        return rows, torch.from_numpy(tab).to(self.base.device)

    def _ft_tables(self, G1, G2, pairs, out):
        """out[l/2][row(A,B)] = packed R2C(g1_{0,A} g1_{l,B} - g2_{0,A} g2_{l,B}) for every pair
        (the ft_ABl maps of bspec.py:591-613); batched over runs of consecutive B at fixed A."""
        b = self.base
        pk = self._packed_layout()
        Nq, N = self.Nk_squeeze, b.N
        runs = []
        for r, (A, B) in enumerate(pairs):
            if runs and runs[-1][0] == A and runs[-1][1] + runs[-1][2] == B and runs[-1][3] + runs[-1][2] == r:
                runs[-1][2] += 1
            else:
                runs.append([A, B, 1, r])
        for li in range(self.Nl):
            for A, B0, nb, r0 in runs:
                gb1 = G1[li * Nq + B0:li * Nq + B0 + nb]
                gb2 = G2[li * Nq + B0:li * Nq + B0 + nb]
                b.r2c(None, out=out[li, r0:r0 + nb], batch=nb,
                      rop=C.rop(C.ROP_DIFF, a=G1[A], sa=0, b=gb1, sb=N, c=G2[A], sc=0, d=gb2, sd=N),
                      kop=C.kop(C.KOP_PACK, pack_index=pk['pack_index'], packed_len=pk['M'], support=self._sup))

    def _q_ainv(self, FT, pair_tab, QA):
        """QA[i] = packed S^-1 P applied to the Q[A^-1 a] map of triplet i (bspec.py:615-666), in chunks of
        triplets of one multipole: the map is assembled from the packed FT rows inside the C2R prologue
        (kop BKQ) and masked / re-transformed by the fused pointing chain; for local l>0 the external-leg
        Legendre factor is the real-space sum_m Y_lm(x) IFFT[Theta_b3 Y_lm(k) FT_0(b1,b2)]
        (apply_real_harmonics, bspec.py:629-636), added as 2l+1 weighted sub-maps.  The number of maps
        per chain comes from the autotuner (knob bk_qmap_chunk), by default as many as fit 12 GiB of
        work maps, at most 64 (32 for chains of several sub-maps)."""
        ab = self.all_bins[self._qperm]  # rows of QA follow this order (sorted by l, then b3)
        if self.chain_streams > 1 and not torch.cuda.is_current_stream_capturing():
            return self._q_ainv_streams(FT, pair_tab, QA, ab, self.chain_streams)
        t0 = 0
        while t0 < self.N_bins:
            ell = int(ab[t0, 3])
            CH = self._qmap_chunk(FT, pair_tab, QA, ell)
            t1 = t0
            while t1 < self.N_bins and t1 - t0 < CH and ab[t1, 3] == ell:
                t1 += 1
            self._q_chunk(FT, pair_tab, QA, t0, t1)
            t0 = t1

    def _q_ainv_streams(self, FT, pair_tab, QA, ab, nstreams):
        """_q_ainv with its chains dealt round-robin to nstreams HIP streams: the current one and side
        streams at the default priority, each with work maps of its own (base.work lanes).  The chains
        are independent (disjoint QA rows), so a VALU-bound row pass of one chain runs beside the
        memory-bound column passes of another.  The first chain runs alone on the current stream, so
        that whatever a chain builds once (tables, integral-constraint maps) exists before the side
        streams start; the current stream waits for every side stream at the end."""
        b = self.base
        chunks, t0 = [], 0
        while t0 < self.N_bins:
            ell = int(ab[t0, 3])
            CH = self._qmap_chunk(FT, pair_tab, QA, ell)  # tuning, if any, runs here on the current stream
            t1 = t0
            while t1 < self.N_bins and t1 - t0 < CH and ab[t1, 3] == ell:
                t1 += 1
            chunks.append((t0, t1))
            t0 = t1
        self._q_chunk(FT, pair_tab, QA, *chunks[0])
        nstreams = self.streams_used = self._streams_that_fit(ab, chunks, nstreams)
        if nstreams < 2:
            for t0, t1 in chunks[1:]:
                self._q_chunk(FT, pair_tab, QA, t0, t1)
            return
        cur = torch.cuda.current_stream(b.device)
        side = getattr(self, '_side_streams', None)
        if side is None or len(side) != nstreams - 1 or side[0].device != b.device:
            side = self._side_streams = [torch.cuda.Stream(device=b.device) for _ in range(nstreams - 1)]
        streams = [cur] + side
        for s in side:
            s.wait_stream(cur)
        try:
            for i, (t0, t1) in enumerate(chunks[1:]):
                lane = (i + 1) % nstreams
                b._work_lane = lane
                with torch.cuda.stream(streams[lane]):
                    self._q_chunk(FT, pair_tab, QA, t0, t1)
        finally:
            b._work_lane = 0
            for s in side:
                cur.wait_stream(s)

    def _streams_that_fit(self, ab, chunks, nstreams):
        """At most nstreams, fewer when the device cannot hold another stream's work maps: each side
        stream keeps the work maps of its largest chain as pb3d_fft_pointing sizes them (per Q map,
        G sub-maps of n0 n1 lz_in complex entries at the chain's input band limit and the output's
        n0 n1 lz_out at k_max), and 2 GiB stay free."""
        b = self.base
        local = b.sightline == 'local'
        n01, nh = int(b.gridsize[0]) * int(b.gridsize[1]), int(b.half_shape[2])
        lz_out = min(nh, self._sup[2] + 1)
        need = 0
        for t0, t1 in chunks:
            G = 2 * int(ab[t0, 3]) + 2 if local and ab[t0, 3] > 0 else 1
            lz_in = min(nh, b.support(self.k_bins_squeeze[int(ab[t0:t1, 2].max()) + 1])[2] + 1)
            need = max(need, 8 * n01 * (t1 - t0) * (G * lz_in + lz_out))
        held = 0  # side streams whose work maps already have that size
        for lane in range(1, nstreams):
            buf = b._bufs.get(('fft_point', lane))
            held += int(buf is not None and buf.numel() >= need)
        free, _ = torch.cuda.mem_get_info(b.device)
        avail = free + torch.cuda.memory_reserved(b.device) - torch.cuda.memory_allocated(b.device)
        fit = held + int(max(0, avail - (2 << 30)) // need)
        return max(1, min(nstreams, 1 + fit))

    def _qmap_chunk(self, FT, pair_tab, QA, ell):
        """Q maps per fused chain for multipole ell: the tuned choice (timed per map on the first maps
        of the multipole, which the caller then recomputes) or the default."""
        b = self.base
        local = b.sightline == 'local'
        ylm = local and ell > 0
        G = 2 * ell + 2 if ylm else 1
        per_map = int(b.gridsize[0]) * int(b.gridsize[1]) * min(b.half_shape[2], self._sup[2] + 1)
        fit = lambda budget: int(max(1, (budget << 30) // ((G + 1) * 8 * per_map)))  # noqa: E731
        # 64 maps per chain for single-sub-map chains; 32 when the harmonic sub-maps multiply the work
        # maps (measured 2 % faster than 54 per chain for the local l = 2 chains at 256^3)
        default = int(max(1, min(64 if G == 1 else 32, fit(12))))
        ab = self.all_bins[self._qperm]
        first = int(np.nonzero(ab[:, 3] == ell)[0][0])
        nl = int(np.sum(ab[:, 3] == ell))
        cands = sorted({c for c in (16, 32, 64, 128, 256) if c <= min(nl, fit(48))} | {default})
        if len(cands) == 1:
            return default
        shape = (tuple(int(v) for v in b.gridsize), self.N_bins, int(self._packed_layout()['M']), ell, G,
                 bool(self.const_mask), getattr(self, '_P', None) is not None and self._P.kind is not None)
        return tune.choose('bk_qmap_chunk', shape, cands,
                           lambda c: self._q_chunk(FT, pair_tab, QA, first, first + min(c, nl)), b.device,
                           default=default, work=lambda c: min(c, nl))

    def _q_chunk(self, FT, pair_tab, QA, t0, t1):
        """The QA rows t0 .. t1 - 1 (one multipole) through one fused chain."""
        b = self.base
        pk = self._packed_layout()
        M, Nq = pk['M'], self.Nk_squeeze
        ab = self.all_bins[self._qperm]
        local = b.sightline == 'local'
        pack = dict(pack_index=pk['pack_index'], packed_len=M, support=self._sup)
        ell = int(ab[t0, 3])
        ylm = local and ell > 0
        nb = t1 - t0
        # before the mask, the Q maps of triplets (b1 <= b2 <= b3) live on modes below edge[b3+1]:
        # the inverse passes are pruned to the chunk's largest b3 (the forward passes keep k_max)
        sup_in = b.support(self.k_bins_squeeze[int(ab[t0:t1, 2].max()) + 1])
        common = dict(bin_table=self.table, pack_index=pk['pack_index'], packed_len=M, trip=self._trip_dev,
                      pair_index=pair_tab, ft=FT, ft_lstride=FT.shape[1] * M, nk=Nq, support=sup_in)
        terms = 2 if ylm else 3
        if self.const_mask:
            # P = mask_mean: the Theta terms need no real-space step (bspec.py:662-664)
            C.check(C.lib().pb3d_kmap_packed(C.ctypes.byref(b.cgrid), None, pk['modes'].data_ptr(), M,
                                             QA[t0:t1].data_ptr(), nb,
                                             C.ctypes.byref(C.kop(C.KOP_BKQ, terms=terms, trip0=t0,
                                                                  leg3_legendre=int(not local),
                                                                  scale=self.mask_mean, **common)),
                                             b.stream()))
            if ylm:
                subs = [(C.kop(C.KOP_BKQ, terms=4, trip0=t0, ell=ell, m=m, **common), None, ell, m, 1.0)
                        for m in range(2 * ell + 1)]
                b.pointing(subs, nb, QA[t0:t1], mask=None, scale=self.mask_mean,
                           kop_out=C.kop(C.KOP_PACK, accumulate=1, **pack))
            return
        subs = [(C.kop(C.KOP_BKQ, terms=terms, trip0=t0, leg3_legendre=int(not local), **common), None,
                 0, 0, 1.0)]
        if ylm:
            subs += [(C.kop(C.KOP_BKQ, terms=4, trip0=t0, ell=ell, m=m, **common), None, ell, m, 1.0)
                     for m in range(2 * ell + 1)]
        # S^-1 P on every Q map (bspec.py:662-666): P of the varying mask, with its integral
        # constraint (the constant-per-bin term subtracted in packed mode space)
        packed = lambda wz: b.r2c(wz, batch=wz.shape[0], kop=C.kop(C.KOP_PACK, **pack),  # noqa: E731
                                  out=torch.empty((wz.shape[0], M), dtype=torch.complex64, device=b.device))
        mask_eff, ic, finish = self._P.chain(False, nb, packed, 'packed')
        b.pointing(subs, nb, QA[t0:t1], mask=mask_eff, scale=1.0, kop_out=C.kop(C.KOP_PACK, **pack), ic=ic)
        finish(QA[t0:t1])

    # ------------------------------------------------------------------ shot noise
    def _bshot(self, gA, gB, gC):
        """assemble_bshot (utils.pyx:641-663): sum over the six assignments of the map sets A, B, C
        to the legs (b1, b2, b3 with its multipole), as triplet sums over one stack [A, B, C]."""
        NG = gA.shape[0]
        st = torch.cat([gA, gB, gC], 0)
        i1, i2, i3 = self._b3_indices()
        n = len(i1)
        sets = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
        j1 = np.concatenate([x * NG + i1 for x, _, _ in sets])
        j2 = np.concatenate([y * NG + i2 for _, y, _ in sets])
        j3 = np.concatenate([z * NG + i3 for _, _, z in sets])
        r = self._triplet(st, j1, j2, j3)
        del st
        return r.reshape(6, n).sum(0)

    def compute_shot_contribution(self, seed, verb=False, data=None, data2=None, cubic_only=False):
        """Shot-noise numerator from one pair of GRFs drawn with 1/P_fid (bspec.py:695-861): the cubic
        term assemble_bshot(g_A, g_B, g_C) / 6 and, unless cubic_only, the quadratic term from the
        singly and doubly weighted data (data, data2 = data by default); returns quadratic - 2 cubic
        (or the cubic term), normalised by V/N/sym."""
        b = self.base
        if not cubic_only:
            assert data is not None, "Must supply data to compute the quadratic shot-noise contribution!"
            if data2 is None:
                print("## Caution: assuming uniform weights!")
                data2 = data
        else:
            print("## Caution: Dropping quadratic (~ P(k)/n) term!")
        assert hasattr(self, 'mask_shot'), "Must supply triply-weighted mask to compute the shot-noise contribution!"
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        if not cubic_only:
            g = self._g_maps(self._filtered_data(data, 'Sinv'))
        if verb:
            print("Computing shot-noise g_{b,l}(r) maps")
        PkA = [b.Pfid[0], 1. / b.Pfid[1]]
        a1 = b.grf_half(seed=seed + int(1e7), Pk_input=PkA, output_type='fourier')
        a2 = b.grf_half(seed=seed + int(2e7), Pk_input=PkA, output_type='fourier')
        A1 = b.applyAinv(a1, invPk0_grid=b.Pk0_grid_host, input_type='fourier', output_type='real')
        A2 = b.applyAinv(a2, invPk0_grid=b.Pk0_grid_host, input_type='fourier', output_type='real')
        Sinv = lambda m, it: b.as_fourier(self.applySinv(m, input_type=it, output_type='fourier'))  # noqa: E731
        gC = self._g_maps(Sinv(A1 * A2 * self.mask_shot, 'real'))
        if not cubic_only:
            d2 = b.as_real(data2)
            gD = self._g_maps(Sinv(A1 * d2, 'real'))
            gE = self._g_maps(Sinv(A2 * d2, 'real'))
        del A1, A2
        gA = self._g_maps(Sinv(a1, 'fourier'))
        gB = self._g_maps(Sinv(a2, 'fourier'))
        del a1, a2
        norm = b.volume / b.gridsize.prod() / self.sym_factor
        if verb:
            print("Assembling cubic shot-noise term")
        cubic = self._bshot(gA, gB, gC) / 6. * norm
        if cubic_only:
            return cubic
        if verb:
            print("Assembling quadratic shot-noise term")
        quad = 0.5 * self._bshot(g, gA, gD) + 0.5 * self._bshot(g, gB, gE)
        quad *= 0.5 * norm
        return quad - 2. * cubic

    def compute_fisher_contribution(self, seed, verb=False):
        """One MC realization (bspec.py:441-445, 473-676): FP64 numpy [N_bins, N_bins], rows indexed
        by the S^-1-weighted triplet, not symmetrised (as the reference)."""
        return self._fisher_contribution_dev(seed).cpu().numpy()

    def _apply_filter_generic(self, a):
        """S^-1 P a with the user's S^-1 (bspec.py:499-505): the mask mean times S^-1 of the Fourier
        map for a constant mask, else S^-1 of the pointing matrix applied to the real map."""
        b = self.base
        if self.const_mask:
            return self.mask_mean * b.as_fourier(self.applySinv(a, input_type='fourier', output_type='fourier'))
        return b.as_fourier(self.applySinv(self.apply_pointing(b.c2r(a)), input_type='real', output_type='fourier'))

    def _q_ainv_generic(self, FT, pair_tab, QA, chunk=8, filt=None):
        """QA rows with the user's S^-1 (bspec.py:615-666): the Q[A^-1 a] maps of a chunk of triplets
        are assembled in real space (C2R with the BKQ prologue; for local l > 0 the external leg's
        sum_m Y_lm(x) IFFT[Theta Y_lm(k) FT_0], bspec.py:629-636), then S^-1 P is the user's callable on
        each map (or filt(real map, 'real'), the covariance filter), and the in-bin modes are packed."""
        b = self.base
        pk = self._packed_layout()
        M, Nq = pk['M'], self.Nk_squeeze
        ab = self.all_bins[self._qperm]
        local = b.sightline == 'local'
        modes = pk['modes'].long()
        common = dict(bin_table=self.table, pack_index=pk['pack_index'], packed_len=M, trip=self._trip_dev,
                      pair_index=pair_tab, ft=FT, ft_lstride=FT.shape[1] * M, nk=Nq, support=self._sup)
        t0 = 0
        while t0 < self.N_bins:
            ell = int(ab[t0, 3])
            ylm = local and ell > 0
            t1 = t0
            while t1 < self.N_bins and t1 - t0 < chunk and ab[t1, 3] == ell:
                t1 += 1
            nb = t1 - t0
            real = b.c2r(None, batch=nb, kop=C.kop(C.KOP_BKQ, terms=2 if ylm else 3, trip0=t0,
                                                   leg3_legendre=int(not local), **common))
            real = real.reshape((nb,) + tuple(int(v) for v in b.gridsize))
            if ylm:
                for m in range(2 * ell + 1):
                    b.c2r(None, out=real, batch=nb, kop=C.kop(C.KOP_BKQ, terms=4, trip0=t0, ell=ell, m=m, **common),
                          rop=C.rop(C.ROP_YLM, ell=ell, m=m, accumulate=1))
            for i in range(nb):
                S = self._apply_filter_generic(b.r2c(real[i])) if filt is None else filt(real[i], 'real')
                QA[t0 + i] = S.reshape(-1)[modes]
            t0 = t1

    def compute_covariance_contribution(self, seed, Pk_cov, verb=False, mask_shot_2pt=None):
        """One realization of the Gaussian covariance for the theory spectrum Pk_cov = [k, P_0 (, P_2, P_4)]
        without shot noise (bspec.py:447-471): the MC-Fisher contribution with S^-1 [P xi P^T + N] S^-dag
        in place of S^-1 P on both sides, N = mask_shot_2pt (default mask_shot).  Every Q map passes
        through the filter one at a time (no fused chain).  A constant mask needs no noise mask."""
        b = self.base
        if not self.const_mask and mask_shot_2pt is None:
            assert getattr(self, 'mask_shot', None) is not None, \
                "Must supply a shot-noise mask to compute the shot-noise contribution!"
            print("## Caution: assuming the same shot-noise mask for the power spectrum and bispectrum!")
            noise = self.mask_shot
        else:
            noise = None if mask_shot_2pt is None else b.as_real(np.asarray(mask_shot_2pt, dtype=np.float64))
        assert len(Pk_cov) != 0, "Must provide fiducial power spectrum if computing covariance!"
        Pk_cov = np.asarray(Pk_cov, dtype=np.float64)
        assert Pk_cov.ndim == 2, "Pk should contain k and P_0 (and optionally P_2, P_4) columns"
        assert len(Pk_cov) in [2, 3, 4], "Pk should contain k and P_0 (and optionally P_2, P_4) columns"
        return self._fisher_contribution_dev(seed, cov=(b.pk_grid(Pk_cov), noise)).cpu().numpy()

    def _contraction_stream(self):
        s = getattr(self, '_cstream', None)
        if s is None or s.device != self.base.device:
            s = self._cstream = torch.cuda.Stream(device=self.base.device)  # default priority
        return s

    def _fisher_contributions_dev(self, seeds, out):
        """out += the Fisher contributions of `seeds` (FP64 device tensor).  With the MFMA contraction,
        each realization's contraction runs on a stream of its own, beside the next realization's GRFs,
        g maps and FT tables; that realization's Q maps (which overwrite the Q-map buffer) wait for it.
        The current stream waits for the last contraction before returning."""
        b = self.base
        seeds = list(seeds)
        if self.precision == 'fp32' or len(seeds) < 2 or torch.cuda.is_current_stream_capturing():
            for sd in seeds:
                self._fisher_contribution_dev(sd, out=out)
            return out
        cur = torch.cuda.current_stream(b.device)
        defer = {'stream': self._contraction_stream(), 'done': None}
        try:
            for sd in seeds:
                self._fisher_contribution_dev(sd, out=out, defer=defer)
        finally:
            cur.wait_stream(defer['stream'])
        return out

    def _fisher_contribution_dev(self, seed, cov=None, out=None, defer=None):
        """One realization's Fisher contribution as an FP64 device tensor [N_bins, N_bins] (added to
        `out` when given).  defer (from _fisher_contributions_dev): run the MFMA contraction on
        defer['stream'] and leave its completion event in defer['done']; the Q maps of the next
        realization wait for that event."""
        b = self.base
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        pk = self._packed_layout()
        M, Nq, NB = pk['M'], self.Nk_squeeze, self.N_bins
        ab = self.all_bins
        if not hasattr(self, '_trip_dev'):
            # Q maps are built in the order (l, b3) so that each chunk has a tight input band limit;
            # the Fisher columns are put back in triplet order after the contraction
            self._qperm = np.lexsort((ab[:, 2], ab[:, 3]))
            self._trip_dev = torch.from_numpy(ab[self._qperm].astype(np.int32).reshape(-1)).to(b.device)
            ainv = set()
            for b1, b2, b3, _ in ab:
                ainv |= {(b1, b2), (b2, b3), (b1, b3)}
            self._ainv_pairs = sorted((int(x), int(y)) for x, y in ainv)
            self._sinv_pairs = sorted(set((int(b1), int(b3)) for b1, _, b3, _ in ab))
            self._ainv_rows, self._ainv_tab = self._pair_rows(self._ainv_pairs)
            self._sinv_rows, _ = self._pair_rows(self._sinv_pairs)
        if getattr(self, '_QA', None) is None:
            self._QA = torch.empty((NB, M), dtype=torch.complex64, device=b.device)
        QA = self._QA
        # GRF pair (bspec.py:648-656)
        a1 = b.grf_half(seed=seed + int(1e7), output_type='fourier')
        a2 = b.grf_half(seed=seed + int(2e7), output_type='fourier')
        # A^-1 side: weighted maps, g maps, FT tables, Q maps (bspec.py:659-666)
        w = b.ainv_weight()
        w1 = b.kmap(a1, C.kop(C.KOP_WEIGHT, w=w)) if w is not None else a1
        w2 = b.kmap(a2, C.kop(C.KOP_WEIGHT, w=w)) if w is not None else a2
        G1, G2 = self._g_maps(w1), self._g_maps(w2)
        FT = torch.empty((self.Nl, len(self._ainv_pairs), M), dtype=torch.complex64, device=b.device)
        self._ft_tables(G1, G2, self._ainv_pairs, FT)
        del G1, G2
        filt = None if cov is None else (lambda m, it: covariance_filter(self, m, it, cov[0], cov[1]))  # noqa: E731
        if defer is not None and defer['done'] is not None:
            torch.cuda.current_stream(b.device).wait_event(defer['done'])  # the previous contraction reads QA
        if filt is not None or self.user_Sinv:
            self._q_ainv_generic(FT, self._ainv_tab, QA, filt=filt)
        else:
            self._q_ainv(FT, self._ainv_tab, QA)
        del FT
        # S^-1 P side (bspec.py:671, compute_Q('Sinv'))
        if filt is not None:
            v1, v2 = filt(a1, 'fourier'), filt(a2, 'fourier')
        elif self.user_Sinv:
            v1, v2 = self._apply_filter_generic(a1), self._apply_filter_generic(a2)
        elif self.const_mask:
            v1 = b.kmap(a1, C.kop(C.KOP_PLAIN, scale=self.mask_mean))
            v2 = b.kmap(a2, C.kop(C.KOP_PLAIN, scale=self.mask_mean))
        else:
            v1 = b.r2c(self.apply_pointing(b.c2r(a1)))
            v2 = b.r2c(self.apply_pointing(b.c2r(a2)))
        H1, H2 = self._g_maps(v1), self._g_maps(v2)
        LT = torch.empty((self.Nl, len(self._sinv_pairs), M), dtype=torch.complex64, device=b.device)
        self._ft_tables(H1, H2, self._sinv_pairs, LT)
        del H1, H2
        # contraction: F[i, j] = sum_{k in b2(i)} degk Re[conj(L_i) QA_j]  (integrate_row, bspec.py:597)
        if self.precision != 'fp32':
            F = torch.zeros((NB, NB), dtype=torch.float64, device=b.device) if out is None else out
            if defer is None:
                self._contract_mfma(LT, QA, F)
                return F
            cur, cs = torch.cuda.current_stream(b.device), defer['stream']
            cs.wait_stream(cur)
            lane = getattr(b, '_work_lane', 0)
            b._work_lane = -1  # the contraction's gather buffer, apart from the chains' work maps
            try:
                with torch.cuda.stream(cs):
                    self._contract_mfma(LT, QA, F)
            finally:
                b._work_lane = lane
            LT.record_stream(cs)
            done = torch.cuda.Event()
            done.record(cs)
            defer['done'] = done
            return F
        F = torch.zeros((NB, NB), dtype=torch.float64, device=b.device)
        QAr = torch.view_as_real(QA)
        nsp = len(self._sinv_pairs)
        for q in range(Nq):
            rows = np.nonzero(ab[:, 1] == q)[0]
            s0, s1 = int(pk['seg'][q]), int(pk['seg'][q + 1])
            if len(rows) == 0 or s1 == s0:
                continue
            src = [(ab[i, 3] // 2) * nsp + self._sinv_rows[(int(ab[i, 0]), int(ab[i, 2]))] for i in rows]
            src_d = torch.from_numpy(np.asarray(src, dtype=np.int32)).to(b.device)
            A = torch.empty((len(rows), s1 - s0), dtype=torch.complex64, device=b.device)
            C.check(C.lib().pb3d_gather_rows(LT.data_ptr(), M, self.Nl * nsp, src_d.data_ptr(), len(rows), s0, s1 - s0,
                                             pk['degk'].data_ptr(), A.data_ptr(), b.stream()))
            Ar = torch.view_as_real(A).reshape(len(rows), 2 * (s1 - s0))
            Br = QAr[:, s0:s1].reshape(NB, 2 * (s1 - s0))
            F[torch.from_numpy(rows).to(b.device)] = torch.matmul(Ar, Br.t()).double()
        # columns of F follow the Q-map order: F[:, qperm[p]] = F_perm[:, p]
        inv = np.empty_like(self._qperm)
        inv[self._qperm] = np.arange(NB)
        F = F[:, torch.from_numpy(inv).to(b.device)]
        sym = torch.from_numpy(self.sym_factor).to(b.device)
        F *= 0.5 * b.volume / b.gridsize.prod() ** 2
        F /= sym[:, None] * sym[None, :]
        if out is not None:
            out += F
            return out
        return F

    def _contract_setup(self):
        """Launch description of the MFMA Fisher contraction, built once per spec: one segment per
        middle bin b2 (its rows: the triangles with that b2, in triplet order; its left rows: their
        S^-1 FT-table rows), the tile list, the row and column maps (columns of the Q-map order go back
        to triplet order in the store) and the 1/sym scales."""
        if getattr(self, '_cplan', None) is not None:
            return self._cplan
        b = self.base
        pk = self._packed_layout()
        ab = self.all_bins
        nsp = len(self._sinv_pairs)
        segs, srcs, rows_all = [], [], []
        a_off = 0
        for q in range(self.Nk_squeeze):
            rows = np.nonzero(ab[:, 1] == q)[0]
            s0, s1 = int(pk['seg'][q]), int(pk['seg'][q + 1])
            if len(rows) == 0 or s1 == s0:
                continue
            src = np.asarray([(ab[i, 3] // 2) * nsp + self._sinv_rows[(int(ab[i, 0]), int(ab[i, 2]))] for i in rows],
                             dtype=np.int32)
            seg = C.FisherSeg()
            seg.a_off, seg.q_off, seg.k, seg.rows, seg.row0 = a_off, 2 * s0, 2 * (s1 - s0), len(rows), len(rows_all)
            segs.append((seg, s0, s1, torch.from_numpy(src).to(b.device)))
            rows_all.extend(rows.tolist())
            a_off += 2 * (s1 - s0) * len(rows)
        arr = (C.FisherSeg * len(segs))(*[x[0] for x in segs])
        nt = C.lib().pb3d_fisher_tiles(arr, len(segs), self.N_bins, None, 0)
        if nt < 0:
            C.check(nt)
        tiles = np.zeros(3 * nt, dtype=np.int32)
        C.check(min(0, C.lib().pb3d_fisher_tiles(arr, len(segs), self.N_bins, tiles.ctypes.data, nt)))
        seg_bytes = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(b.device)  # noqa: E731
        self._cplan = dict(segs=segs, segs_d=dev(seg_bytes), nseg=len(segs), tiles=dev(tiles), ntiles=int(nt),
                           row_map=dev(np.asarray(rows_all, dtype=np.int32)),
                           col_map=dev(self._qperm.astype(np.int32)), left_floats=a_off,
                           scale=dev(1. / self.sym_factor),
                           alpha=0.5 * b.volume / float(b.gridsize.prod()) ** 2)
        return self._cplan

    def _contract_mfma(self, LT, QA, F):
        """F[i, j] += 0.5 V/N^2 / (sym_i sym_j) sum_{k in b2(i)} degk Re[conj(L_i) QA_j] for every triangle
        pair: the left rows of each segment are gathered (degk folded in) into one buffer, then one
        persistent split-BF16 MFMA launch contracts every segment (pb3d_fisher_contract)."""
        b = self.base
        pk = self._packed_layout()
        cp = self._contract_setup()
        M = pk['M']
        left = b.work('fisher_left', 4 * cp['left_floats'])
        nsrc = LT.shape[0] * LT.shape[1]
        for seg, s0, s1, src in cp['segs']:
            C.check(C.lib().pb3d_gather_rows(LT.data_ptr(), M, nsrc, src.data_ptr(), seg.rows, s0, s1 - s0,
                                             pk['degk'].data_ptr(), left.data_ptr() + 4 * seg.a_off, b.stream()))
        prec = self._PREC[self.precision]
        C.check(C.lib().pb3d_fisher_contract(left.data_ptr(), QA.data_ptr(), 2 * M, self.N_bins, cp['segs_d'].data_ptr(),
                                             cp['nseg'], cp['tiles'].data_ptr(), cp['ntiles'],
                                             cp['row_map'].data_ptr(), cp['col_map'].data_ptr(),
                                             cp['scale'].data_ptr(), cp['alpha'], prec, F.data_ptr(), self.N_bins,
                                             b.stream()))

    def compute_fisher(self, N_it, verb=False):
        """Mean over N_it realizations, FP64 on the GPU, then inverted (bspec.py:678-693)."""
        if verb:
            print("Computing %d Fisher contributions" % N_it)
        fish = torch.zeros((self.N_bins, self.N_bins), dtype=torch.float64, device=self.base.device)
        self._fisher_contributions_dev(range(N_it), fish)
        fish /= N_it
        self.fish = fish.cpu().numpy()
        self._inv_fish_dev = torch.linalg.inv(fish)
        self.inv_fish = self._inv_fish_dev.cpu().numpy()
        self._inv_fish_src = self.inv_fish
        return self.fish

    # ------------------------------------------------------------------ unwindowed
    def Bk_unwindowed(self, data, fish=[], include_linear_term=True, verb=False):
        """bspec.py:863-893 (a given Fisher matrix is inverted in FP64 on the GPU, and the inverse is
        applied there)."""
        b = self.base
        if len(fish) != 0:
            self.fish = np.asarray(fish)
            Fd = torch.as_tensor(np.ascontiguousarray(self.fish, dtype=np.float64), device=b.device)
            self._inv_fish_dev = torch.linalg.inv(Fd)
            del Fd
            self.inv_fish = self._inv_fish_dev.cpu().numpy()
            self._inv_fish_src = self.inv_fish
        if not hasattr(self, 'inv_fish'):
            raise Exception("Need to compute Fisher matrix first!")
        num = self.Bk_numerator(data, include_linear_term=include_linear_term, verb=verb)
        # the device copy of inv_fish, re-uploaded if inv_fish was replaced since
        inv = getattr(self, '_inv_fish_dev', None)
        if inv is None or getattr(self, '_inv_fish_src', None) is not self.inv_fish:
            inv = torch.as_tensor(np.ascontiguousarray(self.inv_fish, dtype=np.float64), device=b.device)
            self._inv_fish_dev, self._inv_fish_src = inv, self.inv_fish
        out = (inv @ torch.as_tensor(num, dtype=torch.float64, device=b.device)).cpu().numpy()
        res, idx = {}, 0
        for l in range(0, self.lmax + 1, 2):
            res['b%s' % l] = out[idx:idx + self.N3]
            idx += self.N3
        return res
