"""The pointing matrix of a spatially varying mask on device maps (PolyBin3D/pspec.py:199-213,
PolyBin3D/bspec.py:412-426): a mask multiply, or its global / radial integral-constrained forms
(PolyBin3D/cython/utils.pyx:405-513)

    P   m = mask (m - av_z),    av_z = sum_{x in z} m mask_IC / sum_{x in z} mask_IC,
    P^T m = mask m - mask_IC av'_z,   av'_z = sum_{x in z} m mask / sum_{x in z} mask_IC,

for x in radial bin z (the global constraint: one bin holding every voxel; the radial constraint:
voxels outside every bin map to 0).  Inside the fused pointing chain the bin sums of the combined
real map are taken by the row kernel (pb3d_icsum), and the constant-per-bin term is subtracted in
mode space as sum_z av_z FFT[w 1_z], w = mask (P) or mask_IC (P^T)."""
import numpy as np
import torch

from . import _capi as C

MAX_RADIAL_BINS = 64


class Pointing:
    def __init__(self, base, mask, mask_IC=None, add_GIC=False, add_RIC=False, radial_bins_RIC=()):
        """mask, mask_IC: FP64 host arrays of the grid's shape."""
        b = self.base = base
        self.kind = 'RIC' if add_RIC else ('GIC' if add_GIC else None)
        self.mask = b.as_real(np.ascontiguousarray(mask))
        if self.kind is None:
            return
        mic = np.asarray(mask_IC, dtype=np.float64)
        self.mask_IC = b.as_real(np.ascontiguousarray(mic))
        if self.kind == 'GIC':
            self.nbins, self.rbin, self.rbin_long = 1, None, None
            self.den = np.array([np.sum(mic)])
            self.mask_eff = self.mask
            return
        # radial bins from the FP64 coordinate grid (half-open, utils.pyx:462-463)
        edges = np.asarray(radial_bins_RIC, dtype=np.float64)
        nb = len(edges) - 1
        if not 1 <= nb <= MAX_RADIAL_BINS:
            raise ValueError("radial_bins_RIC: 1 to %d radial bins supported" % MAX_RADIAL_BINS)
        if b.sightline != 'local':
            raise ValueError("the radial integral constraint needs the local sightline's coordinate grid")
        modr = b.modr_grid
        z = np.full(modr.shape, nb, dtype=np.int64)
        for iz in range(nb):
            z[(modr >= edges[iz]) & (modr < edges[iz + 1])] = iz
        self.nbins = nb
        self.den = np.bincount(z.reshape(-1), weights=mic.reshape(-1), minlength=nb + 1)[:nb]
        self.rbin = torch.from_numpy(np.where(z < nb, z, 255).astype(np.uint8)).to(b.device).contiguous()
        self.rbin_long = torch.from_numpy(z.reshape(-1)).to(b.device)
        # the voxels of each radial bin as one segment of a fixed permutation: per-bin sums are segment
        # sums of the permuted map (an atomic histogram into a few bins serialises on the GPU)
        zf = z.reshape(-1)
        order = np.argsort(zf, kind='stable')
        counts = np.bincount(zf, minlength=nb + 1)[:nb]
        self.rperm = torch.from_numpy(order[:int(counts.sum())].astype(np.int64)).to(b.device)
        self.rcounts = torch.from_numpy(counts.astype(np.int64)).to(b.device)
        self.mask_eff = b.as_real(np.ascontiguousarray(mask * (z < nb)))

    # ---------------------------------------------------------------- voxel-space application
    def _bin_means(self, m, w):
        """av_z = sum_{x in z} m w / den_z (FP64) for one real map."""
        mw = (m.double() * w.double()).reshape(-1)
        if self.kind == 'GIC':
            s = mw.sum().reshape(1)
        else:
            s = torch.segment_reduce(mw[self.rperm], 'sum', lengths=self.rcounts)
        return s / self._den_dev()

    def _den_dev(self):
        """The per-bin denominators as an FP64 device tensor, uploaded once."""
        if getattr(self, '_den_t', None) is None:
            self._den_t = torch.from_numpy(np.asarray(self.den, dtype=np.float64)).to(self.base.device)
        return self._den_t

    def _per_voxel(self, av):
        if self.kind == 'GIC':
            return av[0].float()
        ext = torch.cat([av.float(), av.new_zeros(1).float()])
        return ext[self.rbin_long].reshape(self.mask.shape)

    def apply(self, m, transpose=False):
        """P m or P^T m for a device real map (or a batch of them, leading axis)."""
        m = self.base.as_real(m)
        if self.kind is None:
            return m * self.mask
        if m.dim() > 3:
            return torch.stack([self.apply(x, transpose) for x in m])
        if transpose:
            return m * self.mask_eff - self.mask_IC * self._per_voxel(self._bin_means(m, self.mask))
        return (m - self._per_voxel(self._bin_means(m, self.mask_IC))) * self.mask_eff

    # ---------------------------------------------------------------- inside the fused chain
    def chain(self, transpose, batch, corr_maps, tag):
        """For b.pointing: (mask, ic, finish).  mask: the masking map; ic: the bin-sum request;
        finish(out) subtracts sum_z av_z FFT[w 1_z] from the batch of outputs, with corr_maps(w_z) the
        transform of the real maps w_z into the output's format (a [nbins, ...] tensor), computed
        once per (tag, transpose)."""
        if self.kind is None:
            return self.mask, None, lambda out: out
        b = self.base
        weight = self.mask if transpose else self.mask_IC
        key = (tag, transpose)
        if not hasattr(self, '_corr'):
            self._corr = {}
        if key not in self._corr:
            w = self.mask_IC if transpose else self.mask
            if self.kind == 'GIC':
                wz = w.reshape((1,) + tuple(w.shape))
            else:
                wz = torch.stack([w * (self.rbin == z).float() for z in range(self.nbins)])
            self._corr[key] = corr_maps(wz)
        corr = self._corr[key]
        sums = torch.zeros((batch, self.nbins), dtype=torch.float64, device=b.device)
        den = self._den_dev()

        def finish(out):
            coef = (sums / den).to(torch.float32)  # [batch, nbins]
            o = torch.view_as_real(out).reshape(batch, -1)
            c = torch.view_as_real(corr).reshape(self.nbins, -1)
            o -= coef @ c
            return out
        return self.mask_eff, (weight, self.rbin, self.nbins, sums), finish


def covariance_filter(spec, m, input_type, Pk_grid, noise_mask):
    """S^-1 [P xi P^T + N] S^-dag of a device map (real or Fourier, per input_type), the covariance
    counterpart of the S^-1 P filter (PolyBin3D/pspec.py:278-289, 312-325; PolyBin3D/bspec.py:507-519):
    mask_mean^2 xi + mask_mean for a constant mask, else the pointing matrix (with its integral
    constraint) around xi plus the noise mask N.  xi = base.apply_xi of the gridded spectrum Pk_grid
    (base.pk_grid).  Returns a complex64 half-grid map."""
    b = spec.base
    if spec.const_mask:
        Sid = b.as_fourier(spec.applySinv_transpose(m, input_type=input_type, output_type='fourier'))
        CSid = b.apply_xi_half(Sid, None, None, None, Pk_grid, 'fourier')
        CSid.mul_(spec.mask_mean ** 2).add_(Sid, alpha=spec.mask_mean)
        return b.as_fourier(spec.applySinv(CSid, input_type='fourier', output_type='fourier'))
    Sid = b.as_real(spec.applySinv_transpose(m, input_type=input_type, output_type='real'))
    PdSid_real = spec._P.apply(Sid, transpose=True)
    CPdSid = b.apply_xi_half(b.r2c(PdSid_real), PdSid_real, None, None, Pk_grid, 'real')
    Cov = spec._P.apply(CPdSid)
    Cov.addcmul_(Sid, noise_mask)
    return b.as_fourier(spec.applySinv(Cov, input_type='real', output_type='fourier'))
