"""PolyBin3D base class, backend='hip' (MI355X).

Mirrors PolyBin3D.PolyBin3D (PolyBin3D/base.py:6-437): the same constructor arguments, grids,
FFT counters, GRF generator and filtering helpers.  Maps live on the GPU as torch tensors
(float32 real grids [n0,n1,n2], complex64 half grids [n0,n1,n2/2+1]); every transform and map
operation is a hand-written HIP kernel in libpb3d.so reached through ``_capi``.  Host numpy is used
only for set-up tables that must be bit-identical to the reference (FP64 |k| bin membership,
interpolated fiducial spectra) and for the reference's numpy random stream.
"""
import os

import numpy as np
import scipy.interpolate
import torch

from . import _capi as C

PIXEL_WINDOWS = ['none', 'cic', 'tsc', 'pcs', 'interlaced-cic', 'interlaced-tsc', 'interlaced-pcs']


def _pixel_window_1d(kind, x):
    """Voxel windows with first-order alias corrections at x = pi m / n (PolyBin3D/base.py:231-254)."""
    if kind == 'cic':
        s = np.sin(x) ** 2
        return (1. - 2. / 3. * s) ** 0.5
    if kind == 'tsc':
        s = np.sin(x) ** 2
        return (1. - s + 2. / 15 * s ** 2.) ** 0.5
    if kind == 'pcs':
        s = np.sin(x) ** 2
        return (1. - 4. / 3. * s + 2. / 5. * s ** 2 - 4. / 315. * s ** 3) ** 0.5
    if kind == 'interlaced-cic':
        s = np.sinc(x / np.pi)
        v = np.ones_like(s)
        v[s != 0.] = s[s != 0.] ** 2.
        return v
    if kind == 'interlaced-tsc':
        return np.sinc(x / np.pi) ** 3.
    if kind == 'interlaced-pcs':
        return np.sinc(x / np.pi) ** 4.
    raise Exception("Unknown window function '%s' supplied!" % kind)


class YlmTag:
    """Stands for the (2l+1) real spherical harmonics of order l on the real ('real') or Fourier
    ('fourier') grid.  The reference materialises these as arrays (compute_real_harmonics,
    PolyBin3D/cython/utils.pyx:518-609); here the kernels evaluate Y_lm in place, so the tables
    are only labels.  len() gives 2l+1 as the reference arrays do."""

    def __init__(self, ell, space):
        self.ell, self.space = int(ell), space

    def __len__(self):
        return 2 * self.ell + 1


def harmonics_tags(lmax, odd_l, space):
    """Counterpart of compute_real_harmonics(coords, lmax, odd_l) (utils.pyx:518): {l: tag}."""
    return {l: YlmTag(l, space) for l in range(1, lmax + 1) if (l % 2 == 0 or odd_l)}


class PolyBin3D:
    """Base class (PolyBin3D/base.py:6-144) with backend='hip'.

    Inputs as in the reference: boxsize, gridsize, Pk=[k, P0(k)] (None = unit power), boxcenter,
    pixel_window, backend (only 'hip'), nthreads (unused on the GPU), sightline, real_fft (False: the
    reference's complex-grid conventions at the boundary; even n2 only).  Extra: device (default 'cuda'), rng ('numpy' reproduces the reference's legacy
    RandomState stream; 'device' draws on the GPU with the same distribution).
    """

    def __init__(self, boxsize, gridsize, Pk=None, boxcenter=None, pixel_window='none', backend='hip',
                 nthreads=None, sightline='global', real_fft=True, device=None, rng='numpy', verb=False):
        if backend != 'hip':
            raise Exception("polybin3d_amd implements backend='hip' only")
        self.backend = backend
        self.nthreads = nthreads if nthreads is not None else os.cpu_count()
        self.sightline = sightline
        self.pixel_window = pixel_window
        self.real_fft = bool(real_fft)
        self.rng = rng
        self.device = torch.device(device if device is not None else 'cuda')
        if self.device.type != 'cuda':
            raise Exception("backend='hip' needs a GPU device")
        if self.device.index is None:
            self.device = torch.device('cuda', torch.cuda.current_device())
        # base.py:31-60
        self.boxsize = np.asarray([boxsize] * 3 if np.ndim(boxsize) == 0 else boxsize, dtype=np.float64)
        assert len(self.boxsize) == 3, "Need to specify x, y, z boxsize"
        if boxcenter is None:
            self.boxcenter = np.zeros(3)
        else:
            self.boxcenter = np.asarray([boxcenter] * 3 if np.ndim(boxcenter) == 0 else boxcenter, dtype=np.float64)
            assert len(self.boxcenter) == 3, "Need to specify x, y, z boxcenter"
        self.gridsize = np.asarray([gridsize] * 3 if np.ndim(gridsize) == 0 else gridsize, dtype=int)
        assert len(self.gridsize) == 3, "Need to specify n_x, n_y, n_z gridsize"
        if self.real_fft and self.gridsize[2] % 2 != 0:
            raise Exception("Last dimension of grid must be even for real-to-complex FFTs.")
        if not self.real_fft and self.gridsize[2] % 2 != 0:
            # the reference's full-grid frequencies fftshift(arange(-n//2, n//2)) give an odd axis a
            # frequency set that is not symmetric under k -> -k, so Hermitian-paired modes can land in
            # different bins; the half-spectrum transforms here cannot reproduce that
            raise NotImplementedError("backend='hip' with real_fft=False needs an even last grid dimension")
        assert self.sightline in ['local', 'global'], "Sight-line must be 'local' or 'global'"
        if self.sightline == 'global':
            assert np.abs(self.boxcenter).sum() == 0., "Global sightlines should only be used for simulation data!"
        else:
            assert np.abs(self.boxcenter).sum() != 0., "Local sightlines should only be used for lightcone data!"
        # base.py:63-74
        self.volume = self.boxsize.prod()
        self.kF = 2. * np.pi / self.boxsize
        self.kNy = self.gridsize * self.kF * 0.5
        if Pk is None:
            k_tmp = np.arange(0., np.max(self.kNy) * 2, np.min(self.kF) / 10.)
            self.Pfid = np.asarray([k_tmp, 1. + 0. * k_tmp])
        else:
            assert len(Pk) == 2, "Pk should contain k and P_0(k) columns"
            self.Pfid = np.asarray(Pk, dtype=np.float64)
        self.kmin, self.kmax = np.min(self.Pfid[0, 0]), np.max(self.Pfid[0, -1])
        if verb:
            print("# Dimensions: [%.2e, %.2e, %.2e] Mpc/h" % tuple(self.boxsize))
        # Fourier grid (base.py:86-94): fftshift(arange(-n//2, n//2))*kF on axes 0,1, arange on axis 2
        n = self.gridsize
        self.k_arrs = [np.fft.fftshift(np.arange(-n[i] // 2, n[i] // 2)) * self.kF[i] for i in range(2)]
        self.k_arrs += [np.arange(0, n[2] // 2 + 1) * self.kF[2]]
        self.half_shape = (int(n[0]), int(n[1]), int(n[2] // 2 + 1))
        self.full_shape = (int(n[0]), int(n[1]), int(n[2]))
        self.H = int(np.prod(self.half_shape))
        self.N = int(np.prod(n))
        self._modk = None
        self._muk = None
        # line of sight (base.py:111-116)
        if self.sightline == 'global':
            self.sight_vector = np.asarray([0., 0., 1.])
        else:
            self.sight_vector = self.boxcenter / np.sqrt(np.sum(self.boxcenter ** 2))
        assert self.pixel_window in PIXEL_WINDOWS, "Unknown pixel window '%s' supplied!" % pixel_window
        self.cgrid = C.make_grid(n, self.boxsize, self.boxcenter, self.sight_vector, self.sightline == 'local')
        plan = C.c_void_p()
        C.check(C.lib().pb3d_plan_create(C.ctypes.byref(plan), C.ctypes.byref(self.cgrid)))
        self._plan = plan
        # fiducial spectrum on the grid (base.py:122-131), FP64 on host then FP32 on device
        modk = self.modk_half
        Pk0 = scipy.interpolate.interp1d(self.Pfid[0], self.Pfid[1], bounds_error=False, fill_value=0.)(modk)
        invPk0 = np.zeros(modk.shape)
        good = (Pk0 != 0) & np.isfinite(Pk0)
        invPk0[good] = 1. / Pk0[good]
        Pk0[~np.isfinite(Pk0)] = 0.
        invPk0[modk == 0] = 1.
        Pk0[modk == 0] = 1.
        self.Pk0_grid_host = Pk0
        self.invPk0_grid_host = invPk0
        self.invPk0_is_one = bool(np.all(invPk0 == 1.))
        self.invPk0_grid = torch.from_numpy(invPk0.astype(np.float32)).to(self.device)
        if self.pixel_window != 'none':
            w1 = [_pixel_window_1d(self.pixel_window, np.pi / n[i] * self.k_arrs[i] / self.kF[i]) for i in range(3)]
            pw = np.asarray(np.meshgrid(*w1, indexing='ij')).prod(axis=0)
            self.pixel_window_grid_host = pw
            self.pixel_window_grid = torch.from_numpy(pw.astype(np.float32)).to(self.device)
        self.n_FFTs_forward = 0
        self.n_FFTs_reverse = 0
        self._bin_tables = {}
        self._bin_modes = {}
        self._bufs = {}
        self.integrator = IntegrationUtils(self)
        self.map_utils = MapUtils(self)

    def __del__(self):
        try:
            if getattr(self, '_plan', None) is not None and C._lib is not None:
                C.lib().pb3d_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass

    # ------------------------------------------------------------------ host-side grids
    # The transforms work on the half spectrum kz = 0 .. n2/2 whatever real_fft says; with
    # real_fft=False the public k-space grids and Fourier maps take the reference's full-grid form
    # (base.py:86-94: every axis fftshift(arange(-n/2, n/2)) * kF, no degeneracy factor) and are
    # converted at the boundary.
    @property
    def k_grids_half(self):
        return np.meshgrid(*self.k_arrs, indexing='ij')

    @property
    def modk_half(self):
        """|k| on the half grid in FP64 exactly as the reference computes it (base.py:91-92)."""
        if self._modk is None:
            kg = self.k_grids_half
            self._modk = np.sqrt(kg[0] ** 2 + kg[1] ** 2 + kg[2] ** 2)
        return self._modk

    @property
    def muk_half(self):
        if self._muk is None:
            self._muk = self._mu_of(self.k_grids_half, self.modk_half)
        return self._muk

    def _mu_of(self, kg, modk):
        mu = np.zeros(modk.shape)
        nz = modk != 0
        mu[nz] = np.sum(self.sight_vector[:, None] * np.asarray(kg)[:, nz], axis=0) / modk[nz]
        return mu

    @property
    def k_grids(self):
        if self.real_fft:
            return self.k_grids_half
        n = self.gridsize
        return np.meshgrid(*[np.fft.fftshift(np.arange(-n[i] // 2, n[i] // 2)) * self.kF[i] for i in range(3)],
                           indexing='ij')

    @property
    def modk_grid(self):
        if self.real_fft:
            return self.modk_half
        kg = self.k_grids
        return np.sqrt(kg[0] ** 2 + kg[1] ** 2 + kg[2] ** 2)

    @property
    def muk_grid(self):
        if self.real_fft:
            return self.muk_half
        kg = self.k_grids
        return self._mu_of(kg, np.sqrt(kg[0] ** 2 + kg[1] ** 2 + kg[2] ** 2))

    @property
    def degeneracy_factor(self):
        """1 + (kz > 0) on the half grid (base.py:93-96); None with real_fft=False."""
        if not self.real_fft:
            return None
        return np.asarray(1. + (self.k_grids_half[2] > 0), dtype=np.float64)

    def to_full(self, x, parity=1):
        """Half-spectrum map(s) [..., n0, n1, n2/2+1] -> the full grid [..., n0, n1, n2] with
        X(-k) = parity * conj X(k): Hermitian (parity 1, the transform of a real map) or anti-Hermitian
        (parity -1)."""
        n0, n1, n2 = self.full_shape
        h = n2 // 2 + 1
        out = torch.empty(tuple(x.shape[:-3]) + self.full_shape, dtype=x.dtype, device=x.device)
        out[..., :h] = x
        # X[i, j, k] = parity * conj X[-i, -j, n2 - k] for k > n2/2
        src = x[..., 1:n2 - h + 1].conj()
        src = torch.roll(torch.flip(src, dims=(-3, -2, -1)), shifts=(1, 1), dims=(-3, -2))
        out[..., h:] = src if parity > 0 else -src
        return out

    def full_parts(self, x):
        """Full-grid map(s) -> the half spectra (kz = 0 .. n2/2) of their Hermitian and anti-Hermitian
        parts, (X(k) + conj X(-k)) / 2 and (X(k) - conj X(-k)) / 2."""
        h = self.full_shape[2] // 2 + 1
        neg = torch.roll(torch.flip(x, dims=(-3, -2, -1)), shifts=(1, 1, 1), dims=(-3, -2, -1))[..., :h].conj()
        return (0.5 * (x[..., :h] + neg)).contiguous(), (0.5 * (x[..., :h] - neg)).contiguous()

    def from_full(self, x):
        """Full-grid map(s) -> the half spectrum of their Hermitian part: the map whose C2R is
        Re IFFT(X)."""
        return self.full_parts(x)[0]

    def fourier_linear(self, x, f, parity):
        """f (a k-space map on half spectra that multiplies mode k by a factor g(k) with
        g(-k) = parity * conj g(k)) applied to a Fourier map in the caller's form.  A full grid
        (real_fft=False) is split into its Hermitian and anti-Hermitian parts, each mapped on the half
        spectrum and completed with its own symmetry, so that the result is f on every full-grid mode."""
        x = x.to(device=self.device, dtype=torch.complex64) if torch.is_tensor(x) else x
        if self.real_fft or not (tuple(np.shape(x)[-3:]) == self.full_shape):
            return self.public_fourier(f(self.as_fourier(x)))
        if not torch.is_tensor(x):
            x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex64)).to(self.device)
        hp, ap = self.full_parts(x.contiguous())
        return self.to_full(f(hp), parity) + self.to_full(f(ap), -parity)

    def public_fourier(self, x):
        """A Fourier map as the caller sees it: the half spectrum, or the full grid with real_fft=False."""
        return x if self.real_fft else self.to_full(x)

    def wrap_user_filter(self, f):
        """A user S^-1 (applySinv / applySinv_transpose) as the reference calls it: with real_fft=False its
        Fourier inputs are full grids (its outputs go through as_fourier, which takes either form)."""
        if f is None or self.real_fft:
            return f

        def g(m, input_type='real', output_type='real'):
            if input_type == 'fourier' and torch.is_tensor(m) and tuple(m.shape[-3:]) == self.half_shape:
                m = self.to_full(m)
            return f(m, input_type=input_type, output_type=output_type)
        return g

    @property
    def r_grids(self):
        n = self.gridsize
        r = [np.fft.fftshift(np.arange(-n[i] // 2, n[i] // 2)) * self.boxsize[i] / n[i] + self.boxcenter[i]
             for i in range(3)]
        return np.meshgrid(*r, indexing='ij')

    @property
    def modr_grid(self):
        """|r| of every voxel in FP64 (the attribute pspec.py:110 and bspec.py:135 set for the radial
        integral constraint and the wide-angle terms), computed once."""
        if getattr(self, '_modr', None) is None:
            r = self.r_grids
            self._modr = np.sqrt(r[0] ** 2. + r[1] ** 2. + r[2] ** 2.)
        return self._modr

    def bin_index_host(self, edges):
        """int64 host array: b with edges[b] <= |k| < edges[b+1] in FP64 (utils.pyx:63), else -1."""
        edges = np.asarray(edges, dtype=np.float64)
        nb = len(edges) - 1
        if not np.all(np.diff(edges) > 0):
            raise ValueError("bin edges must be strictly increasing")
        idx = np.searchsorted(edges, self.modk_half, side='right') - 1
        idx[(idx < 0) | (idx >= nb)] = -1
        return idx

    def bin_table(self, edges):
        """uint8 device table: index b with edges[b] <= |k| < edges[b+1] (FP64, as
        utils.pyx:63), or 255.  Cached per edge list."""
        edges = np.asarray(edges, dtype=np.float64)
        key = tuple(edges.tolist())
        if key not in self._bin_tables:
            nb = len(edges) - 1
            if nb < 1 or nb > 254:
                raise ValueError("between 1 and 254 bins supported")
            if not np.all(np.diff(edges) > 0):
                raise ValueError("bin edges must be strictly increasing")
            idx = np.searchsorted(edges, self.modk_half, side='right') - 1
            idx[(idx < 0) | (idx >= nb)] = C.PB3D_NO_BIN
            self._bin_tables[key] = torch.from_numpy(idx.astype(np.uint8)).to(self.device)
            # the in-bin modes (ascending linear index), for reductions that visit only them
            inb = np.flatnonzero(idx.reshape(-1) != C.PB3D_NO_BIN).astype(np.int32)
            self._bin_modes[key] = torch.from_numpy(inb).to(self.device)
        return self._bin_tables[key]

    def bin_modes(self, edges):
        """int32 device list of the modes inside the bins of `edges` (ascending linear index)."""
        self.bin_table(edges)
        return self._bin_modes[tuple(np.asarray(edges, dtype=np.float64).tolist())]

    # ------------------------------------------------------------------ buffers
    def real_zeros(self, size=None, numpy=False):
        shape = self._shape(size, tuple(int(v) for v in self.gridsize))
        return np.zeros(shape) if numpy else torch.zeros(shape, dtype=torch.float32, device=self.device)

    def complex_zeros(self, size=None, numpy=False):
        shape = self._shape(size, self.half_shape)
        return np.zeros(shape, dtype=np.complex128) if numpy else torch.zeros(shape, dtype=torch.complex64,
                                                                                 device=self.device)

    @staticmethod
    def _shape(size, base):
        if size is None:
            return base
        if np.ndim(size) == 0:
            return (int(size),) + base
        return tuple(int(s) for s in size) + base

    def work(self, name, nbytes):
        """Reusable device scratch (stream-ordered use only).  Work that runs on a side stream sets
        _work_lane to that stream's index, so that it gets buffers of its own."""
        lane = getattr(self, '_work_lane', 0)
        key = name if not lane else (name, lane)
        b = self._bufs.get(key)
        if b is None or b.numel() < nbytes:
            b = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=self.device)
            self._bufs[key] = b
        return b

    def stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _input_extent(self, kop, batch):
        """Complex elements a k-space input of `batch` entries spans: one H-map per
        kop.x_batch_group entries, x_batch_stride apart."""
        if kop is None:
            return batch * self.H
        return ((batch - 1) // max(1, int(kop.x_batch_group))) * int(kop.x_batch_stride) + self.H

    def _need(self, t, n, what):
        """Raise unless tensor t is a contiguous device tensor holding at least n elements of its
        kind (complex64 modes or float32 voxels), so no kernel reads or writes past it."""
        if t is None:
            raise ValueError("%s: missing tensor" % what)
        if not t.is_contiguous() or t.device != self.device:
            raise ValueError("%s: tensor must be contiguous on %s" % (what, self.device))
        if t.numel() < n:
            raise ValueError("%s: tensor has %d elements, needs %d" % (what, t.numel(), n))

    def _check_rop(self, rop, batch, what):
        """Operands of a voxel op must hold (batch-1)*stride + N voxels."""
        if rop is None:
            return
        for name in ('a', 'b', 'c', 'd'):
            t = getattr(rop, '_t', {}).get(name)
            if t is not None:
                self._need(t, (batch - 1) * getattr(rop, 'batch_stride_' + name) + self.N, what + ' rop.' + name)

    def _check_kop(self, kop, batch, what):
        """Tables of a mode op must cover the half grid (bins: one entry per batch entry)."""
        if kop is None:
            return
        t = getattr(kop, '_t', {})
        for name in ('bin_table', 'w', 'pack_index'):
            if name in t:
                self._need(t[name], self.H, what + ' kop.' + name)
        if 'bins' in t:
            self._need(t['bins'], batch, what + ' kop.bins')
        if kop.kind == C.KOP_BKQ:
            self._need(t.get('trip'), 4 * (kop.trip0 + batch), what + ' kop.trip')
            self._need(t.get('pair_index'), kop.nk * kop.nk, what + ' kop.pair_index')
            self._need(t.get('ft'), kop.packed_len, what + ' kop.ft')

    # ------------------------------------------------------------------ data movement
    def as_real(self, x):
        """float32 contiguous device tensor of shape [...,n0,n1,n2]."""
        if isinstance(x, np.ndarray):
            x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
        if not torch.is_tensor(x):
            raise TypeError("expected a numpy array or torch tensor")
        x = x.to(device=self.device, dtype=torch.float32).contiguous()
        if tuple(x.shape[-3:]) != tuple(int(v) for v in self.gridsize):
            raise ValueError("real map has shape %s, grid is %s" % (tuple(x.shape), tuple(self.gridsize)))
        return x

    def as_fourier(self, x):
        """complex64 contiguous device tensor of shape [...,n0,n1,n2/2+1]."""
        if isinstance(x, np.ndarray):
            x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex64))
        if not torch.is_tensor(x):
            raise TypeError("expected a numpy array or torch tensor")
        x = x.to(device=self.device, dtype=torch.complex64).contiguous()
        if not self.real_fft and tuple(x.shape[-3:]) == self.full_shape:
            return self.from_full(x)
        if tuple(x.shape[-3:]) != self.half_shape:
            raise ValueError("Fourier map has shape %s, half grid is %s" % (tuple(x.shape), self.half_shape))
        return x

    # ------------------------------------------------------------------ FFTs
    def r2c(self, x, out=None, rop=None, kop=None, batch=None, accumulate=False):
        """Batched fused R2C.  x: real tensor (or None when rop.a is set)."""
        if batch is None:
            batch = 1 if x is None or x.dim() == 3 else int(np.prod(x.shape[:-3]))
        if out is None:
            out = torch.empty((batch,) + self.half_shape if batch > 1 else self.half_shape,
                              dtype=torch.complex64, device=self.device)
        needs_work = kop is not None and (bool(kop.accumulate) or kop.kind == C.KOP_PACK)
        work = self.work('fft_r2c', batch * self.H * 8) if needs_work else None
        if x is not None:
            self._need(x, batch * self.N, 'r2c input')
        elif rop is None or not rop.a:
            raise ValueError("r2c: no input")
        self._check_rop(rop, batch, 'r2c')
        self._check_kop(kop, batch, 'r2c')
        packed = kop is not None and kop.kind == C.KOP_PACK
        self._need(out, batch * (kop.packed_len if packed else self.H), 'r2c output')
        C.check(C.lib().pb3d_fft_r2c(self._plan, C._ptr(x), out.data_ptr(), C._ptr(work), batch,
                                     C.byref_or_none(rop), C.byref_or_none(kop), self.stream()))
        self.n_FFTs_forward += batch
        return out

    def c2r(self, k, out=None, kop=None, rop=None, batch=None):
        """Batched fused C2R with 1/N (k may be None for the BKQ prologue)."""
        if batch is None:
            batch = 1 if k.dim() == 3 else int(np.prod(k.shape[:-3]))
        if out is None:
            g = tuple(int(v) for v in self.gridsize)
            out = torch.empty((batch,) + g if batch > 1 else g, dtype=torch.float32, device=self.device)
        work = self.work('fft_c2r', batch * self.H * 8)
        if k is not None:
            self._need(k, self._input_extent(kop, batch), 'c2r input')
        elif kop is None or kop.kind != C.KOP_BKQ:
            raise ValueError("c2r: no input")
        self._check_kop(kop, batch, 'c2r')
        self._check_rop(rop, batch, 'c2r')
        self._need(out, batch * self.N, 'c2r output')
        C.check(C.lib().pb3d_fft_c2r(self._plan, C._ptr(k), out.data_ptr(), work.data_ptr(), batch,
                                     C.byref_or_none(kop), C.byref_or_none(rop), self.stream()))
        self.n_FFTs_reverse += batch
        return out

    def to_fourier(self, _input_rmap):
        """Unnormalised R2C (base.py:176-200); float32 maps on the GPU.  With real_fft=False the full
        complex grid (numpy fftn of the real map)."""
        x = self.as_real(_input_rmap)
        return self.public_fourier(self.r2c(x))

    def to_real(self, _input_kmap):
        """C2R with 1/N (base.py:202-229); a full grid (real_fft=False) gives Re IFFT, as the reference."""
        k = self.as_fourier(_input_kmap)
        return self.c2r(k)

    def support(self, kmax):
        """Frequency-index radii (per axis) covering every mode with |k| < kmax, for band-limited
        (pruned) transforms: |fs(i)| * kF_i <= |k| < kmax."""
        return [int(min(np.floor(kmax / self.kF[i]), self.gridsize[i] // 2)) for i in range(3)]

    def mask_rows(self, mask, weight=None):
        """Row flags for pb3d_fft_pointing: uint8 per (x, y) row, 0 on the aligned groups of 64 rows
        where the mask (and the integral-constraint weight) is zero on every voxel, whose outputs are
        then zero rows written without transforms; None when every group has a nonzero voxel.  Built
        once per mask (and weight) tensor and in-place version."""
        cache = self._bufs.setdefault('mask_rows', [])
        wv = None if weight is None else weight._version
        for m, mv, w, wv_, f in cache:
            if m is mask and mv == mask._version and w is weight and wv_ == wv:
                return f
        n0, n1, n2 = (int(v) for v in self.gridsize)
        rows, group = n0 * n1, 64
        on = (mask.reshape(rows, n2) != 0).any(dim=1)
        if weight is not None:
            on |= (weight.reshape(rows, n2) != 0).any(dim=1)
        ng = (rows + group - 1) // group
        grp = torch.nn.functional.pad(on, (0, ng * group - rows)).reshape(ng, group).any(dim=1)
        f = None if bool(grp.all()) else grp.repeat_interleave(group)[:rows].to(torch.uint8).contiguous()
        cache.append((mask, mask._version, weight, wv, f))
        del cache[:-16]  # every (mask, weight) pair of a few specs on this base
        return f

    def pointing(self, subs, batch, out, mask=None, scale=1.0, kop_out=None, ic=None):
        """out[b] = kop_out(R2C(mask * scale * sum_g coef_g Y_g(x) C2R(kop_g(in_g))[b])), fused
        (pb3d_fft_pointing).  subs: list of (kop, input or None for BKQ, ell, m, coef).  ic: optional
        (weight, bin or None, nbins, sums): sums[b, z] += sum over bin z of m_b * weight, m_b the
        combined real map before the mask (integral constraints)."""
        G = len(subs)
        arr = (C.SubMap * G)()
        keep = []
        for g, (kop, inp, ell, m, coef) in enumerate(subs):
            self._check_kop(kop, batch, 'pointing sub %d' % g)
            if inp is not None:
                self._need(inp, self._input_extent(kop, batch), 'pointing sub %d input' % g)
            elif kop.kind != C.KOP_BKQ:
                raise ValueError("pointing: sub-map %d has no input" % g)
            arr[g].kop = kop
            arr[g].inp = C._ptr(inp)
            arr[g].ell, arr[g].m, arr[g].coef = int(ell), int(m), float(coef)
            keep.append((kop, inp))
        if mask is not None:
            self._need(mask, self.N, 'pointing mask')
        self._check_kop(kop_out, batch, 'pointing output')
        packed = kop_out is not None and kop_out.kind == C.KOP_PACK
        self._need(out, batch * (kop_out.packed_len if packed else self.H), 'pointing output')
        nw = C.lib().pb3d_fft_pointing_work_size(C.ctypes.byref(self.cgrid), batch, G, arr, C.byref_or_none(kop_out))
        if nw < 0:
            C.check(-1)
        work = self.work('fft_point', nw * 8)
        icd = None
        if ic is not None:
            weight, rbin, nbins, sums = ic
            if weight.dtype != torch.float32 or sums.dtype != torch.float64:
                raise ValueError("pointing ic: float32 weight and float64 sums expected")
            self._need(weight, self.N, 'pointing ic weight')
            if rbin is not None and (rbin.dtype != torch.uint8 or rbin.numel() < self.N):
                raise ValueError("pointing ic bin: uint8 device tensor of the grid's size expected")
            self._need(sums, batch * nbins, 'pointing ic sums')
            npart = C.lib().pb3d_icsum_partial_size(C.ctypes.byref(self.cgrid), batch, int(nbins))
            if npart < 0:
                C.check(-1)
            part = self.work('icsum_partial', 8 * npart)
            icd = C.IcSum(weight.data_ptr(), C._ptr(rbin), int(nbins), 0, sums.data_ptr(), part.data_ptr())
            keep.append((weight, rbin, sums))
        rows = None if mask is None else self.mask_rows(mask, None if ic is None else ic[0])
        C.check(C.lib().pb3d_fft_pointing(self._plan, batch, G, arr, C._ptr(mask), float(scale),
                                          C.byref_or_none(kop_out), C.byref_or_none(icd), C._ptr(rows),
                                          out.data_ptr(), work.data_ptr(), self.stream()))
        self.n_FFTs_reverse += batch * G
        self.n_FFTs_forward += batch
        return out

    def kmap(self, x, kop, out=None, batch=1):
        if out is None:
            out = torch.empty((batch,) + self.half_shape if batch > 1 else self.half_shape,
                              dtype=torch.complex64, device=self.device)
        self._need(x, self._input_extent(kop, batch), 'kmap input')
        self._need(out, batch * self.H, 'kmap output')
        self._check_kop(kop, batch, 'kmap')
        C.check(C.lib().pb3d_kmap(C.ctypes.byref(self.cgrid), x.data_ptr(), out.data_ptr(), batch,
                                  C.ctypes.byref(kop), self.stream()))
        return out

    # ------------------------------------------------------------------ GRFs
    def _Pk_grid_host(self, Pk_input):
        if len(Pk_input) == 0:
            return self.Pk0_grid_host
        assert len(np.asarray(Pk_input).shape) == 2, "Pk should contain k and P_0, (and optionally P_2, P_4) columns"
        assert len(Pk_input) in [2, 3, 4], "Pk should contain k and P_0, (and optionally P_2, P_4) columns"
        modk = self.modk_half
        P = scipy.interpolate.interp1d(Pk_input[0], Pk_input[1], bounds_error=False, fill_value=0.)(modk)
        mu = self.muk_half
        if len(Pk_input) > 2:
            P = P + 0.5 * (3 * mu ** 2 - 1) * scipy.interpolate.interp1d(
                Pk_input[0], Pk_input[2], bounds_error=False, fill_value=0.)(modk)
        if len(Pk_input) > 3:
            P = P + (3 - 30 * mu ** 2 + 35 * mu ** 4) / 8. * scipy.interpolate.interp1d(
                Pk_input[0], Pk_input[3], bounds_error=False, fill_value=0.)(modk)
        return P

    def generate_data(self, seed=None, Pk_input=[], output_type='real', include_pixel_window=True):
        """Gaussian random field (base.py:305-353).  rng='numpy': the reference's legacy
        RandomState(seed).randn stream, drawn on the host and symmetrised on the GPU.  Fourier output:
        the half spectrum, or the full Hermitian grid with real_fft=False."""
        out = self.grf_half(seed, Pk_input, output_type, include_pixel_window)
        return self.public_fourier(out) if output_type == 'fourier' else out

    def grf_half(self, seed=None, Pk_input=[], output_type='real', include_pixel_window=True):
        """generate_data with a half-spectrum Fourier output (what the estimators use internally)."""
        assert output_type in ['fourier', 'real'], "Valid output types are 'fourier' and 'real' only!"
        key = ('sqrtPw', tuple(map(tuple, np.asarray(Pk_input))) if len(Pk_input) else None,
               bool(include_pixel_window and self.pixel_window != 'none'))
        if key not in self._bufs:
            P = self._Pk_grid_host(Pk_input)
            with np.errstate(invalid='ignore'):
                s = np.sqrt(P)
            if self.pixel_window != 'none' and include_pixel_window:
                s = s * self.pixel_window_grid_host
            self._bufs[key] = torch.from_numpy(s.astype(np.float32)).to(self.device)
        sqrtPw = self._bufs[key]
        n = tuple(int(v) for v in self.gridsize)
        out = torch.empty(self.half_shape, dtype=torch.complex64, device=self.device)
        norm = float(self.N / np.sqrt(self.volume))
        if self.rng == 'device':
            s = int(seed) if seed is not None else int(np.random.randint(0, 2 ** 62))
            seed_dev = getattr(self, '_seed_dev', None)
            if seed_dev is not None:
                # inside a graph capture: seed = (device seed base) + s, set before every replay
                C.check(C.lib().pb3d_grf_device_at(C.ctypes.byref(self.cgrid), seed_dev.data_ptr(), s,
                                                   sqrtPw.data_ptr(), norm, out.data_ptr(), self.stream()))
            else:
                C.check(C.lib().pb3d_grf_device(C.ctypes.byref(self.cgrid), s, sqrtPw.data_ptr(), norm,
                                                out.data_ptr(), self.stream()))
        else:
            rs = np.random.RandomState(seed) if seed is not None else np.random.mtrand._rand
            re = torch.from_numpy(rs.randn(*n)).to(self.device)
            im = torch.from_numpy(rs.randn(*n)).to(self.device)
            C.check(C.lib().pb3d_grf_hermitian(C.ctypes.byref(self.cgrid), re.data_ptr(), im.data_ptr(),
                                               sqrtPw.data_ptr(), norm, out.data_ptr(), self.stream()))
        if output_type == 'fourier':
            return out
        return self.c2r(out)

    # ------------------------------------------------------------------ filters
    def applySinv_trivial(self, input_data, input_type='real', output_type='real'):
        """base.py:355-362."""
        if input_type == output_type:
            return input_data
        if input_type == 'fourier':
            return self.to_real(input_data)
        return self.to_fourier(input_data)  # the caller's Fourier form (full grid with real_fft=False)

    def ainv_weight(self, invPk0_grid=None):
        """invPk0 / W^2 as a device FP32 half grid, or None when it is identically one
        (applyAinv, base.py:364-389)."""
        if invPk0_grid is None and self.pixel_window == 'none':
            return None if self.invPk0_is_one else self.invPk0_grid
        # cached per array: the entry keeps the array it was made from, so that an id reused by another
        # (temporary) array is never served a stale table; a few arrays are kept
        cache = self._bufs.setdefault('ainv_weights', [])
        for arr, t in cache:
            if arr is invPk0_grid:
                return t
        inv = self.invPk0_grid_host if invPk0_grid is None else np.asarray(invPk0_grid)
        if not self.real_fft and inv.shape == self.full_shape:
            inv = inv[..., :self.half_shape[2]]  # a weight built on the full-grid |k|, even under k -> -k
        w = inv.copy()
        if self.pixel_window != 'none':
            pw = self.pixel_window_grid_host
            nz = pw != 0
            w[nz] = w[nz] / pw[nz] ** 2
            w[~nz] = 0.
        t = torch.from_numpy(w.astype(np.float32)).to(self.device)
        cache.append((invPk0_grid, t))
        del cache[:-4]
        return t

    def applyAinv(self, input_data, invPk0_grid=None, input_type='real', output_type='real'):
        """base.py:364-389: multiply by invPk0 and divide by the pixel window twice."""
        w = self.ainv_weight(invPk0_grid)
        op = C.kop(C.KOP_WEIGHT, w=w) if w is not None else None
        if input_type == 'real':
            f = self.r2c(self.as_real(input_data), kop=op)
        elif output_type == 'fourier':
            # the weight is even under k -> -k; a full grid keeps its anti-Hermitian part too
            return self.fourier_linear(input_data, lambda x: self.kmap(x, op) if op is not None else x.clone(), 1)
        else:
            f = self.as_fourier(input_data)
            f = self.kmap(f, op) if op is not None else f.clone()
        return self.c2r(f) if output_type == 'real' else self.public_fourier(f)

    def apply_fourier_harmonics(self, a_map_real, Ylm_real, Ylm_fourier, out=None):
        """sum_m FFT[a(x) Y_lm(x)] Y_lm(k) (base.py:402-411), one fused R2C per m."""
        ell = Ylm_real.ell if isinstance(Ylm_real, YlmTag) else int(Ylm_real)
        a = self.as_real(a_map_real)
        if out is None:
            out = torch.empty(self.half_shape, dtype=torch.complex64, device=self.device)
        for m in range(2 * ell + 1):
            self.r2c(a, out=out, rop=C.rop(C.ROP_YLM, ell=ell, m=m),
                     kop=C.kop(C.KOP_YLM, ell=ell, m=m, accumulate=m > 0))
        return out

    def apply_real_harmonics(self, a_map_fourier, Ylm_real, Ylm_fourier, out=None):
        """sum_m IFFT[a(k) Y_lm(k)] Y_lm(x) (base.py:391-400), one fused C2R per m."""
        ell = Ylm_real.ell if isinstance(Ylm_real, YlmTag) else int(Ylm_real)
        a = self.as_fourier(a_map_fourier)
        if out is None:
            out = torch.empty(tuple(int(v) for v in self.gridsize), dtype=torch.float32, device=self.device)
        for m in range(2 * ell + 1):
            self.c2r(a, out=out, kop=C.kop(C.KOP_YLM, ell=ell, m=m, x_batch_stride=0),
                     rop=C.rop(C.ROP_YLM, ell=ell, m=m, accumulate=m > 0))
        return out

    def pk_grid(self, Pk_cov):
        """The gridded theory spectrum of apply_xi (pspec.py:237-251) as FP32 device half grids:
        global {0: P_0 + L_2(mu) P_2 + L_4(mu) P_4}, local {0: P_0, 2: P_2, 4: P_4} (columns present
        in Pk_cov = [k, P_0 (, P_2, P_4)]), interpolated in FP64 with zero outside the table."""
        Pk_cov = np.asarray(Pk_cov, dtype=np.float64)
        modk = self.modk_half
        interp = lambda col: scipy.interpolate.interp1d(Pk_cov[0], Pk_cov[col], bounds_error=False,  # noqa: E731
                                                        fill_value=0.)(modk)
        if self.sightline == 'global':
            P = interp(1)
            mu = self.muk_half
            if len(Pk_cov) > 2:
                P = P + 0.5 * (3 * mu ** 2 - 1) * interp(2)
            if len(Pk_cov) > 3:
                P = P + (3 - 30 * mu ** 2 + 35 * mu ** 4) / 8. * interp(3)
            grids = {0: P}
        else:
            grids = {0: interp(1)}
            for col, l in ((2, 2), (3, 4)):
                if len(Pk_cov) > col:
                    grids[l] = interp(col)
        return {l: torch.from_numpy(g.astype(np.float32)).to(self.device) for l, g in grids.items()}

    def _grid_weight(self, g):
        """A half-grid weight map (numpy array, as the reference's callers build them on modk_grid, or a
        tensor) as an FP32 device tensor.  A numpy grid is uploaded once and reused while the same
        array is passed (a few grids are kept)."""
        if torch.is_tensor(g):
            t = g.to(device=self.device, dtype=torch.float32).contiguous()
        else:
            cache = self._bufs.setdefault('grid_weights', [])
            for arr, t in cache:
                if arr is g:
                    break
            else:
                t = torch.from_numpy(np.ascontiguousarray(g, dtype=np.float32)).to(self.device)
                cache.append((g, t))
                del cache[:-4]
        if not self.real_fft and tuple(t.shape) == self.full_shape:
            # a spectrum built on the full-grid modk / muk: its half-spectrum part (the kz = -n2/2 plane
            # has the |k| and mu^2 of kz = +n2/2)
            t = t[..., :self.half_shape[2]].contiguous()
        if tuple(t.shape) != self.half_shape:
            raise ValueError("a gridded spectrum must have the half-grid shape %s" % (self.half_shape,))
        return t

    def apply_xi(self, input_map_fourier, input_map_real=None, Ylm_real=None, Ylm_fourier=None, Pk_grid={},
                 output_type='real'):
        """apply_xi_half with the caller's Fourier form (the full grid with real_fft=False)."""
        out = self.apply_xi_half(input_map_fourier, input_map_real, Ylm_real, Ylm_fourier, Pk_grid, output_type)
        return self.public_fourier(out) if output_type == 'fourier' else out

    def apply_xi_half(self, input_map_fourier, input_map_real=None, Ylm_real=None, Ylm_fourier=None, Pk_grid={},
                      output_type='real'):
        """xi(x, y) applied symmetrically for a gridded spectrum (base.py:413-437), Pk_grid a dict of half
        grids (numpy, as the reference's callers build them, or device tensors, e.g. from pk_grid):
        P_0 a + sum_l [P_l sum_m Y_lm(k) FFT[Y_lm a_real] + sum_m Y_lm(x) IFFT[Y_lm(k) P_l a]] / 2 on
        the local sightline, P(k, mu) a on the global one.  The weights ride in the fused transforms'
        mode ops.  Without input_map_real the local multipoles use the C2R of the input (the reference
        requires it)."""
        a = self.as_fourier(input_map_fourier)
        Pk_grid = {l: self._grid_weight(g) for l, g in Pk_grid.items()}
        P = self.kmap(a, C.kop(C.KOP_WEIGHT, w=Pk_grid[0]))
        ells = [l for l in Pk_grid if l != 0]
        if self.sightline == 'global' or not ells:
            return P if output_type == 'fourier' else self.c2r(P)
        ar = self.c2r(a) if input_map_real is None else self.as_real(input_map_real)
        P_real = torch.empty(tuple(int(v) for v in self.gridsize), dtype=torch.float32, device=self.device)
        h = torch.empty(self.half_shape, dtype=torch.complex64, device=self.device)
        first = True
        for l in ells:
            self.apply_fourier_harmonics(ar, l, l, out=h)
            self.kmap(h, C.kop(C.KOP_WEIGHT, w=Pk_grid[l], scale=0.5, accumulate=1), out=P)
            for m in range(2 * l + 1):
                self.c2r(a, out=P_real, kop=C.kop(C.KOP_YLM, ell=l, m=m, w=Pk_grid[l], x_batch_stride=0),
                         rop=C.rop(C.ROP_YLM, ell=l, m=m, accumulate=not first, scale=0.5))
                first = False
        if output_type == 'fourier':
            return self.r2c(P_real, out=P, kop=C.kop(C.KOP_PLAIN, accumulate=1))
        out = self.c2r(P)
        out += P_real
        return out

    # ------------------------------------------------------------------ reductions
    def bin_reduce(self, x, edges, ells, y=None, ells2=None, real=False, y_ostride=None, y_qstride=0,
                   nq=None):
        """out[q][o][b] = sum_{k in b} degk Re[x_q conj(y_{q,o})] L_ells[o] L_ells2[o]  (FP64 tensor)."""
        table = self.bin_table(edges)
        nb = len(edges) - 1
        nout = len(ells)
        if nq is None:
            nq = 1 if x.dim() == 3 else int(np.prod(x.shape[:-3]))
        xs = self.H
        ws = C.lib().pb3d_bin_reduce_work_size(C.ctypes.byref(self.cgrid), nq, nout, nb)
        work = self.work('bins', ws)
        out = torch.empty((nq, nout, nb), dtype=torch.float64, device=self.device)
        l1 = C.int_array(ells)
        l2 = C.int_array(ells2 if ells2 is not None else [0] * nout)
        if y_ostride is None:
            y_ostride = self.H
        # every read stays inside the tensors handed in
        self._need(x, nq * self.H if not real else nq * self.H, 'bin_reduce x')
        if y is not None:
            self._need(y, (nq - 1) * y_qstride + (nout - 1) * y_ostride + self.H, 'bin_reduce y')
        modes = self.bin_modes(edges)
        C.check(C.lib().pb3d_bin_reduce_modes(C.ctypes.byref(self.cgrid), x.data_ptr(), xs, C._ptr(y), y_qstride,
                                              y_ostride, nq, nout, l1, l2, 1 if real else 0, table.data_ptr(), nb,
                                              modes.data_ptr(), int(modes.numel()), out.data_ptr(),
                                              work.data_ptr(), self.stream()))
        return out


class IntegrationUtils:
    """The reference's IntegrationUtils contract (PolyBin3D/cython/utils.pyx:17-179) on GPU maps:
    one fused pass over the half grid per call instead of one sweep per bin."""

    def __init__(self, base):
        self.base = base

    def _parts(self, spec):
        """(Hermitian, anti-Hermitian) half spectra of a full grid (real_fft=False), else (spec, None)."""
        b = self.base
        if not b.real_fft and tuple(np.shape(spec)[-3:]) == b.full_shape:
            x = spec if torch.is_tensor(spec) else torch.from_numpy(np.ascontiguousarray(spec, dtype=np.complex64))
            return b.full_parts(x.to(device=b.device, dtype=torch.complex64).contiguous())
        return b.as_fourier(spec), None

    def _pairs(self, spec1, spec2, l):
        """Half-spectrum pairs whose degk-weighted sums of Re[x conj y] L_l add up to the full-grid sum of
        Re[s1 conj s2] L_l: with H, A the Hermitian and anti-Hermitian parts, Re[H1 conj H2] and
        Re[A1 conj A2] are even under k -> -k and the cross terms odd, so an even L_l keeps the former
        and an odd L_l the latter (the half spectrum of a real_fft=True map is its own H)."""
        h1, a1 = self._parts(spec1)
        h2, a2 = self._parts(spec2)
        if a1 is None and a2 is None:
            return [(h1, h2)]
        pairs = [(h1, h2), (a1, a2)] if l % 2 == 0 else [(h1, a2), (a1, h2)]
        return [(x, y) for x, y in pairs if x is not None and y is not None]

    def cross_integrate(self, spec1, spec2, k_edges, l):
        """sum_{k in bin} degk Re[s1 conj s2] L_l(mu) (utils.pyx:51-78)."""
        b = self.base
        out = np.zeros(len(k_edges) - 1)
        for x, y in self._pairs(spec1, spec2, l):
            out = out + b.bin_reduce(x, k_edges, [l], y=y)[0, 0].cpu().numpy()
        return out

    def cross_integrate_wcoeff(self, coeff, spec1, spec2, k_edges):
        """sum_{k in bin} degk Re[coeff_bin s1 conj s2], one complex coefficient per bin
        (utils.pyx:139-157): Re(coeff) sum Re[s1 conj s2] - Im(coeff) sum Im[s1 conj s2], the second sum
        being sum Re[s1 conj(i s2)]."""
        b = self.base
        c = np.asarray(coeff, dtype=np.complex128)
        s2 = spec2 if torch.is_tensor(spec2) else torch.from_numpy(np.ascontiguousarray(spec2, dtype=np.complex64))
        s2 = s2.to(device=b.device, dtype=torch.complex64)
        re = self.cross_integrate(spec1, s2, k_edges, 0)
        im = self.cross_integrate(spec1, (1j * s2).contiguous(), k_edges, 0)
        return c.real * re - c.imag * im

    def integrate(self, spec, k_edges, l):
        """sum_{k in bin} degk |s|^2 L_l(mu) (utils.pyx:80-107)."""
        b = self.base
        if b.real_fft or tuple(np.shape(spec)[-3:]) != b.full_shape:
            return b.bin_reduce(b.as_fourier(spec), k_edges, [l])[0, 0].cpu().numpy()
        return self.cross_integrate(spec, spec, k_edges, l)

    def real_integrate_double(self, spec, l1, l2, k_edges):
        b = self.base
        s = spec if torch.is_tensor(spec) else torch.from_numpy(np.asarray(spec, dtype=np.float32))
        s = s.to(device=b.device, dtype=torch.float32).contiguous()
        if not b.real_fft and tuple(s.shape[-3:]) == b.full_shape:
            s = s[..., :b.half_shape[2]].contiguous()  # a real spectrum symmetric under k -> -k
        return b.bin_reduce(s, k_edges, [l1], ells2=[l2], real=True)[0, 0].cpu().numpy()

    def integrate_row(self, input_map, input_array, kmin, kmax):
        """out[j] = sum_{kmin<=k<kmax} degk Re[conj(m) Q_j] for every map Q_j (utils.pyx:162-179)."""
        b = self.base
        out = 0.
        for Q, m in self._pairs(input_array, input_map, 0):
            nq = Q.shape[0]
            out = out + b.bin_reduce(Q, [kmin, kmax], [0], y=m, nq=nq, y_qstride=0)[:, 0, 0].cpu().numpy()
        return out


class MapUtils:
    """The reference's MapUtils contract (PolyBin3D/cython/utils.pyx:184-383) on GPU maps."""

    def __init__(self, base):
        self.base = base

    def fourier_filter(self, kmap, l, kmin, kmax):
        b = self.base
        op = C.kop(C.KOP_FILTER, bin_table=b.bin_table([kmin, kmax]), ell=l)
        return b.fourier_linear(kmap, lambda x: b.kmap(x, op), (-1) ** l)

    def fourier_filter_ll(self, kmap, l1, l2, kmin, kmax):
        b = self.base
        op = C.kop(C.KOP_FILTER, bin_table=b.bin_table([kmin, kmax]), ell=l1, ell2=l2)
        return b.fourier_linear(kmap, lambda x: b.kmap(x, op), (-1) ** (l1 + l2))

    def _weight(self, map2):
        """A real k-space weight as a device FP32 half grid (a full grid with real_fft=False gives its
        half: the callers' weights are even under k -> -k)."""
        b = self.base
        w = map2 if torch.is_tensor(map2) else torch.from_numpy(np.asarray(map2, dtype=np.float32))
        w = w.to(device=b.device, dtype=torch.float32).contiguous()
        if not b.real_fft and tuple(w.shape[-3:]) == b.full_shape:
            w = w[..., :b.half_shape[2]].contiguous()
        return w

    def prod_fourier(self, map1, map2):
        """map1 times a real k-space weight that is even under k -> -k (as the reference's callers'
        |k|- and mu^2-dependent grids are) (utils.pyx:213-228)."""
        b = self.base
        op = C.kop(C.KOP_WEIGHT, w=self._weight(map2))
        return b.fourier_linear(map1, lambda x: b.kmap(x, op), 1)

    def div_fourier(self, map1, map2):
        """map1 / map2 for a real k-space map2, zero where map2 is zero (utils.pyx:230-243): the
        weight 1 / map2 is formed once on the device and applied in the mode op."""
        b = self.base
        w = self._weight(map2)
        inv = torch.where(w != 0, 1. / torch.where(w != 0, w, torch.ones_like(w)), torch.zeros_like(w))
        op = C.kop(C.KOP_WEIGHT, w=inv.contiguous())
        return b.fourier_linear(map1, lambda x: b.kmap(x, op), 1)

    def prod_fourier_filter(self, map1, map2, kmin, kmax):
        """map1 map2 on the modes kmin <= |k| < kmax, zero elsewhere (utils.pyx:325-338)."""
        b = self.base
        wop = C.kop(C.KOP_WEIGHT, w=self._weight(map2))
        fop = C.kop(C.KOP_FILTER, bin_table=b.bin_table([kmin, kmax]), ell=0)
        return b.fourier_linear(map1, lambda x: b.kmap(b.kmap(x, wop), fop), 1)

    def prod_real(self, map1, map2):
        """map1 map2 in real space (utils.pyx:248-260)."""
        return self.base.as_real(map1) * self.base.as_real(map2)

    def div_real(self, map1, map2):
        """map1 / map2 in real space, zero where map2 is zero (utils.pyx:265-278)."""
        a, d = self.base.as_real(map1), self.base.as_real(map2)
        return torch.where(d != 0, a / torch.where(d != 0, d, torch.ones_like(d)), torch.zeros_like(a))

    def prod_real_diff(self, map1, map2, map3, map4):
        """map1 map2 - map3 map4 in real space (utils.pyx:371-383)."""
        r = self.base.as_real
        return r(map1) * r(map2) - r(map3) * r(map4)

    def sum_pair(self, map1, map2):
        """sum_x map1 map2, accumulated in FP64 (utils.pyx:343-352)."""
        r = self.base.as_real
        return float(torch.sum(r(map1).double() * r(map2).double()).item())

    def sum_triplet(self, map1, map2, map3):
        """sum_x map1 map2 map3, accumulated in FP64 (utils.pyx:357-366)."""
        r = self.base.as_real
        return float(torch.sum(r(map1).double() * r(map2).double() * r(map3).double()).item())
