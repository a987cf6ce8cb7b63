"""ctypes bindings of ``libcdft.so`` (C ABI declared in ``include/cdft.h``)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

CDFT_OK, CDFT_ETA_TOO_LARGE, CDFT_UPHILL_LIMIT = 0, 1, 2
CDFT_BAD_ARG, CDFT_HIP_ERROR, CDFT_NOT_SUPPORTED = -1, -2, -3
MODEL_NONE, ROSENFELD, RSLT, ESFMT, WHITEBEAR_I, WHITEBEAR_II = -1, 0, 1, 2, 3, 4
FP64, MIXED = 0, 1
WEIGHTS_ACCURATE, WEIGHTS_REFERENCE = 0, 1  # cdft_weights_mode
AO_FORCE_EXACT, AO_FORCE_REFERENCE = 0, 1  # cdft_ao_force
DEVICE_PTRS, ACCUMULATE, FMT_ONLY, NO_IDEAL, SQUARED_MATRIX, DYNAMIC_MATRIX, CENTRAL_DIFFERENCES = 1, 2, 4, 8, 16, 32, 64
FIXED_BACKGROUND, HOMOGENEOUS_BOUNDARY = 1, 2  # enum cdft_boundary

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p


class CdftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cdft error {code}: {msg}")
        self.code = code


class EtaTooLarge(CdftError):
    """Mirror of the reference's Eta_Too_Large_Exception (include/FMT.h:649-663)."""


class _Desc(ctypes.Structure):
    _fields_ = [
        ("Nx", _i), ("Ny", _i), ("Nz", _i),
        ("dx", _d), ("dy", _d), ("dz", _d),
        ("model", _i), ("A", _d), ("B", _d), ("hsd", _d),
        ("precision", _i),
        ("fmt_weights", _p),
        ("ideal", _i), ("mu", _d), ("vext", _p),
        ("mean_field", _i), ("lj_sigma", _d), ("lj_eps", _d), ("lj_rcut", _d), ("kT", _d),
        ("mf_kernel", _p),
        ("mf_potential", _i), ("mf_alpha", _d), ("mf_bh", _i), ("mf_r_att_min", _d), ("mf_scheme", _i),
        ("mf_ngauss", _i),
        ("ao_Rp", _d), ("ao_reservoir", _d), ("ao_weights", _p),
        ("fixed_mass", _d), ("boundary", _i), ("mf_scale", _d), ("weights_mode", _i), ("ao_force", _i),
    ]


class _MixDesc(ctypes.Structure):
    _fields_ = [("Nx", _i), ("Ny", _i), ("Nz", _i), ("dx", _d), ("dy", _d), ("dz", _d), ("model", _i), ("A", _d),
                ("B", _d), ("nspecies", _i), ("hsd", _p), ("fmt_weights", _p), ("weights_mode", _i), ("ideal", _i),
                ("mu", _p), ("vext", _p), ("fixed_mass", _p), ("boundary", _p)]


class _PotDesc(ctypes.Structure):
    _fields_ = [("kind", _i), ("sigma", _d), ("eps", _d), ("rcut", _d), ("alpha", _d), ("bh", _i),
                ("r_att_min", _d)]


# mean-field potentials and kernel rules (cdft_potential, cdft_mf_scheme)
POT_LJ, POT_TWF, POT_WHDF = 0, 1, 2
MF_QF, MF_ZERO, MF_LE, MF_QE, MF_LF, MF_GAUSS_E, MF_GAUSS_F = range(7)


class FireParams(ctypes.Structure):
    _fields_ = [("dt", _d), ("dt_max", _d), ("dt_min", _d), ("alpha_start", _d), ("f_dec", _d), ("f_inc", _d),
                ("f_alf", _d), ("N_delay", _i), ("N_P_negative_max", _i), ("initial_delay", _i)]


class FireReport(ctypes.Structure):
    _fields_ = [("F", _d), ("fmax_alias", _d), ("fmax_raw", _d), ("dt", _d), ("alpha", _d), ("vv", _d),
                ("rms_force", _d), ("monitor", _d), ("steps", _i), ("calls", _i)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OpenDdftReport(ctypes.Structure):
    _fields_ = [("F", _d), ("dt", _d), ("time", _d), ("rhs_max", _d), ("calls", _i), ("restarts", _i),
                ("evaluations", _i)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DdftReport(ctypes.Structure):
    _fields_ = [("F", _d), ("dt", _d), ("time", _d), ("rhs_max", _d), ("mass", _d), ("calls", _i), ("restarts", _i)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def library_path():
    return os.path.join(_HERE, "libcdft.so")


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python -m classicaldft_amd.build`")
        L = ctypes.CDLL(path)
        L.cdft_last_error.restype = ctypes.c_char_p
        L.cdft_device_count.argtypes = [ctypes.POINTER(_i)]
        L.cdft_system_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(_p)]
        L.cdft_system_destroy.argtypes = [_p]
        L.cdft_system_eval.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_fmt_eval.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_mixture_create.argtypes = [ctypes.POINTER(_MixDesc), ctypes.POINTER(_p)]
        L.cdft_mixture_destroy.argtypes = [_p]
        L.cdft_mixture_species.argtypes = [_p, ctypes.POINTER(_i)]
        L.cdft_mixture_fmt_eval.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_mixture_eval.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_mixture_chem_potential.argtypes = [_p, _p]
        L.cdft_mixture_fire_create.argtypes = [_p, _p, _i, ctypes.POINTER(FireParams), ctypes.POINTER(_p)]
        L.cdft_system_stream.argtypes = [_p, ctypes.POINTER(_p)]
        L.cdft_system_set_stream.argtypes = [_p, _p]
        L.cdft_set_device.argtypes = [_i]
        L.cdft_get_device.argtypes = [ctypes.POINTER(_i)]
        L.cdft_system_a_vdw.argtypes = [_p, ctypes.POINTER(_d)]
        L.cdft_system_layout.argtypes = [_p, ctypes.POINTER(_i), ctypes.POINTER(_i)]
        L.cdft_system_get_weights.argtypes = [_p, _p]
        L.cdft_system_get_mf_kernel.argtypes = [_p, _p]
        L.cdft_lj_hsd.argtypes = [_d, _d, _d, _d, ctypes.POINTER(_d)]
        L.cdft_system_batch_create.argtypes = [ctypes.POINTER(_Desc), _i, _p, _p, ctypes.POINTER(_p)]
        L.cdft_system_autotune.argtypes = [_p, _i, ctypes.POINTER(_d)]
        L.cdft_fire_batch_create.argtypes = [ctypes.POINTER(_Desc), _i, _p, _p, _p, ctypes.POINTER(FireParams),
                                             ctypes.POINTER(_p)]
        L.cdft_fire_batch_run.argtypes = [_p, _i, _d, _i]
        L.cdft_fire_batch_set_precision.argtypes = [_p, _i]
        L.cdft_fire_batch_status.argtypes = [_p, _p, _p, ctypes.POINTER(_i)]
        L.cdft_fire_batch_get.argtypes = [_p, _i, _p]
        L.cdft_fire_batch_state.argtypes = [_p, _i, _p, _p, _p, _p]
        L.cdft_fire_batch_system.argtypes = [_p, ctypes.POINTER(_p)]
        L.cdft_fire_batch_destroy.argtypes = [_p]
        L.cdft_fire_batch_destroy.restype = None
        L.cdft_fire_group_create.argtypes = [_p, _i, ctypes.POINTER(_p)]
        L.cdft_fire_group_run.argtypes = [_p, _i, _d]
        L.cdft_fire_group_destroy.argtypes = [_p]
        L.cdft_fire_group_destroy.restype = None
        L.cdft_system_hessian_dot.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_system_dynamic_dot.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_system_hessian_dot_noalloc.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_system_eigen_objective.argtypes = [_p, _p, _p, _p, _d, _i, _i, ctypes.POINTER(_d), _p]
        L.cdft_system_log_det.argtypes = [_p, _p, _i, _d, _d, ctypes.c_long, _i, _i, _i, ctypes.POINTER(_d),
                                          ctypes.POINTER(_d)]
        L.cdft_potential_hsd.argtypes = [ctypes.POINTER(_PotDesc), _d, ctypes.POINTER(_d)]
        L.cdft_potential_vdw.argtypes = [ctypes.POINTER(_PotDesc), _d, ctypes.POINTER(_d)]
        L.cdft_potential_watt.argtypes = [ctypes.POINTER(_PotDesc), _i, _p, _p]
        L.cdft_fmt_hessian_dot.argtypes = [_p, _p, _p, _p, _i]
        L.cdft_unet_create.argtypes = [_i, ctypes.c_ulonglong, ctypes.POINTER(_p)]
        L.cdft_unet_destroy.argtypes = [_p]
        L.cdft_unet_destroy.restype = None
        L.cdft_unet_num_params.argtypes = [_p, ctypes.POINTER(ctypes.c_longlong)]
        L.cdft_unet_get_params.argtypes = [_p, _p]
        L.cdft_unet_set_params.argtypes = [_p, _p]
        L.cdft_unet_forward.argtypes = [_p, _p, _p, _p, _i, _p]
        L.cdft_unet_grad.argtypes = [_p, _p, _p, _p, _p, _i, ctypes.POINTER(ctypes.c_float), _p]
        L.cdft_unet_train_step.argtypes = [_p, _p, _p, _p, _p, _i, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
        L.cdft_fft_r2c.argtypes = [_i, _i, _i, _i, _p, _p]
        L.cdft_fft_c2r.argtypes = [_i, _i, _i, _i, _p, _p]
        L.cdft_system_eval_async.argtypes = [_p, _p, _p, _i]
        L.cdft_system_fetch.argtypes = [_p, _p]
        L.cdft_system_time.argtypes = [_p, _p, _p, _i, _i, ctypes.POINTER(_d)]
        L.cdft_malloc.argtypes = [ctypes.POINTER(_p), ctypes.c_ulonglong]
        L.cdft_free.argtypes = [_p]
        L.cdft_memcpy_h2d.argtypes = [_p, _p, ctypes.c_ulonglong]
        L.cdft_memcpy_d2h.argtypes = [_p, _p, ctypes.c_ulonglong]
        L.cdft_system_bulk.argtypes = [_p, _d, ctypes.POINTER(_d), ctypes.POINTER(_d)]
        L.cdft_system_chem_potential.argtypes = [_p, _p]
        L.cdft_system_set_fixed_masses.argtypes = [_p, _p]
        L.cdft_fire_default_params.argtypes = [ctypes.POINTER(FireParams)]
        L.cdft_fire_create.argtypes = [_p, _p, _i, ctypes.POINTER(FireParams), ctypes.POINTER(_p)]
        L.cdft_fire_step.argtypes = [_p, ctypes.POINTER(FireReport)]
        L.cdft_fire_run.argtypes = [_p, _i, _d, ctypes.POINTER(FireReport)]
        L.cdft_fire_get.argtypes = [_p, _p, _p, _p, _p]
        L.cdft_fire_set_fixed_direction.argtypes = [_p, _p, _i]
        L.cdft_fire_reset.argtypes = [_p, _p, _i]
        L.cdft_fire_destroy.argtypes = [_p]
        L.cdft_ddft_open_create.argtypes = [_p, _p, _d, _d, _d, _d, ctypes.POINTER(_p)]
        L.cdft_ddft_open_step.argtypes = [_p, ctypes.POINTER(OpenDdftReport)]
        L.cdft_ddft_open_get_density.argtypes = [_p, _p]
        L.cdft_ddft_open_destroy.argtypes = [_p]
        L.cdft_dst1_3d.argtypes = [_i, _i, _i, _p]
        L.cdft_mf_kernel_points.argtypes = [_i, _i, _i, _d, _d, _d, ctypes.POINTER(_PotDesc), _p, _i, _i, _p]
        L.cdft_ddft_create.argtypes = [_p, _p, _d, _d, _d, ctypes.POINTER(_p)]
        L.cdft_ddft_step.argtypes = [_p, ctypes.POINTER(DdftReport)]
        L.cdft_ddft_set_max_time_step.argtypes = [_p, _d]
        L.cdft_ddft_set_central_differences.argtypes = [_p, _i]
        L.cdft_ddft_get_density.argtypes = [_p, _p]
        L.cdft_ddft_destroy.argtypes = [_p]
        L.cdft_system_set_timing.argtypes = [_p, _i]
        L.cdft_system_get_timing.argtypes = [_p, ctypes.POINTER(_i), _p, ctypes.POINTER(ctypes.c_char_p)]
        _LIB = L
    return _LIB


class DdftStalled(CdftError):
    """CDFT_DDFT_STALLED: the DDFT fixed point fails at the smallest time step the reference can set."""


def _check(rc):
    if rc == CDFT_OK:
        return
    msg = lib().cdft_last_error().decode(errors="replace")
    if rc == CDFT_ETA_TOO_LARGE:
        raise EtaTooLarge(rc, msg)
    if rc == 3:
        raise DdftStalled(rc, msg)
    raise CdftError(rc, msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _out(a, n, shape):
    """A caller-supplied output array the library writes n doubles into: float64, C-contiguous, n elements."""
    if a is None:
        return np.zeros(shape)
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous or not a.flags.writeable:
        raise ValueError("output array must be a writeable, C-contiguous float64 numpy array")
    if a.size != n:
        raise ValueError(f"output array must hold {n} values, got {a.size}")
    return a


def device_count():
    n = _i(0)
    _check(lib().cdft_device_count(ctypes.byref(n)))
    return n.value


def set_device(device):
    """cdft_set_device: the device on which this thread creates its systems."""
    _check(lib().cdft_set_device(int(device)))


def get_device():
    n = _i(0)
    _check(lib().cdft_get_device(ctypes.byref(n)))
    return n.value


class System:
    """One grid + FMT (+ ideal gas, LJ mean field, V_ext, mu) on the GPU."""

    def __init__(self, shape, dx, hsd=1.0, model=WHITEBEAR_I, precision=FP64, weights=None, ideal=False, mu=0.0,
                 vext=None, mean_field=False, sigma=1.0, eps=1.0, rcut=2.5, kT=1.0, mf_kernel=None, A=1.0, B=0.0,
                 mus=None, vexts=None, potential=POT_LJ, alpha=0.0, bh=False, r_att_min=0.0, mf_scheme=MF_QF,
                 ngauss=0, ao_Rp=0.0, ao_reservoir=0.0, ao_weights=None, fixed_mass=0.0, fixed_background=False,
                 homogeneous_boundary=False, mf_scale=1.0, weights_mode=WEIGHTS_ACCURATE,
                 ao_force=AO_FORCE_EXACT):
        """mus (and optionally vexts, shape (nb,) + shape) make a batch of nb members (cdft_system_batch_create).
        potential / alpha / bh / r_att_min / mf_scheme / ngauss choose the mean-field potential and kernel rule;
        ao_Rp > 0 adds an Asakura-Oosawa polymer of that radius at reservoir density ao_reservoir.
        fixed_mass / fixed_background / homogeneous_boundary: Species::setFixedMass / setFixedBackground /
        setHomogeneousBoundary (the full evaluation rescales rho to the mass and constrains dF); batches take one
        fixed mass per member.  weights_mode: how generated cell weights are computed (WEIGHTS_ACCURATE on the
        device, or WEIGHTS_REFERENCE: the reference's double expressions on the host).  ao_force: AO_FORCE_EXACT
        (the derivative of the AO free energy) or AO_FORCE_REFERENCE (each off-diagonal tensor measure counted
        once, as FMT_AO_Species::setFundamentalMeasures_AO does, include/Species.h:406-425)."""
        fm = np.atleast_1d(np.asarray(fixed_mass, dtype=np.float64)).reshape(-1)
        self.shape = tuple(int(n) for n in shape)
        self.fixed_background = bool(fixed_background)
        dxs = (dx, dx, dx) if np.isscalar(dx) else tuple(dx)
        self.dx = dxs
        self.dV = dxs[0] * dxs[1] * dxs[2]
        self.N = int(np.prod(self.shape))
        keep = []

        def arr(a, n):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            if a.size != n:
                raise ValueError(f"expected {n} values, got {a.size}")
            keep.append(a)
            return a.ctypes.data

        d = _Desc(self.shape[0], self.shape[1], self.shape[2], dxs[0], dxs[1], dxs[2], int(model), float(A),
                  float(B), float(hsd), int(precision), arr(weights, 11 * self.N), int(bool(ideal)), float(mu),
                  arr(vext, self.N), int(bool(mean_field)), float(sigma), float(eps), float(rcut), float(kT),
                  arr(mf_kernel, self.N), int(potential), float(alpha), int(bool(bh)), float(r_att_min),
                  int(mf_scheme), int(ngauss), float(ao_Rp), float(ao_reservoir), arr(ao_weights, 11 * self.N),
                  float(fm[0]), (1 if fixed_background else 0) | (2 if homogeneous_boundary else 0), float(mf_scale),
                  int(weights_mode), int(ao_force))
        h = _p()
        if mus is None:
            self.nb = 1
            _check(lib().cdft_system_create(ctypes.byref(d), ctypes.byref(h)))
        else:
            m = np.ascontiguousarray(mus, dtype=np.float64).reshape(-1)
            self.nb = m.size
            vx = arr(vexts, self.nb * self.N)
            _check(lib().cdft_system_batch_create(ctypes.byref(d), self.nb, m.ctypes.data_as(_p), _p(vx),
                                                  ctypes.byref(h)))
        self.h = h
        self.model = model
        self.precision = precision
        if fm.size > 1:
            if fm.size != self.nb:
                raise ValueError(f"expected {self.nb} fixed masses, got {fm.size}")
            _check(lib().cdft_system_set_fixed_masses(self.h, _ptr(np.ascontiguousarray(fm))))

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- evaluations on host arrays --
    def _ashape(self):
        return self.shape if self.nb == 1 else (self.nb,) + self.shape

    def eval(self, rho, dF=None, accumulate=False, only_fex=False):
        """DFT::calculateFreeEnergyAndDerivatives(only_fex): returns (parts[5], dF); batches: parts (nb, 5).
        only_fex leaves out the ideal-gas term (the external field and mu stay, as in the reference)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.size != self.nb * self.N:
            raise ValueError(f"expected {self.nb * self.N} densities, got {rho.size}")
        out = _out(dF, self.nb * self.N, self._ashape())
        parts = np.zeros(5 * self.nb)
        flags = (ACCUMULATE if accumulate else 0) | (NO_IDEAL if only_fex else 0)
        _check(lib().cdft_system_eval(self.h, _ptr(rho), _ptr(out), _ptr(parts), flags))
        return (parts if self.nb == 1 else parts.reshape(self.nb, 5)), out

    def fmt_eval(self, rho, dF=None, accumulate=False):
        """FMT::calculateFreeEnergyAndDerivatives: returns (F_hs, dF); batches: F_hs (nb,)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.size != self.nb * self.N:
            raise ValueError(f"expected {self.nb * self.N} densities, got {rho.size}")
        out = _out(dF, self.nb * self.N, self._ashape())
        F = np.zeros(self.nb)
        _check(lib().cdft_fmt_eval(self.h, _ptr(rho), _ptr(out), _ptr(F), ACCUMULATE if accumulate else 0))
        return (F[0] if self.nb == 1 else F), out

    def hessian_dot(self, rho, v, ideal=True, out=None, accumulate=False):
        """DFT::matrix_dot_v_intern: (d2 beta Omega / drho drho) v (ideal = full_hessian_)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if rho.size != self.nb * self.N or v.size != rho.size:
            raise ValueError(f"expected {self.nb * self.N} densities and directions")
        res = _out(out, self.nb * self.N, self._ashape())
        fl = (0 if ideal else NO_IDEAL) | (ACCUMULATE if accumulate else 0)
        _check(lib().cdft_system_hessian_dot(self.h, _ptr(rho), _ptr(v), _ptr(res), fl))
        return res


    def dynamic_dot(self, rho, v, ideal=True, squared=False, central=False):
        """DDFT::matrix_dot_v_intern (src/DDFT.cpp:353-361): g . (H v), g the DDFT mobility operator (central:
        the DDFT's central_differences)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if rho.size != self.N or v.size != self.N:
            raise ValueError(f"expected {self.N} densities and components")
        out = np.zeros(rho.shape)
        fl = (0 if ideal else NO_IDEAL) | (SQUARED_MATRIX if squared else 0) | (CENTRAL_DIFFERENCES if central else 0)
        _check(lib().cdft_system_dynamic_dot(self.h, _ptr(rho), _ptr(v), _ptr(out), fl))
        return out

    def eigen_objective(self, rho, x, vshift=None, scale=1.0, change_sign=False, ideal=True, squared=False,
                        dynamic=False):
        """(f, df) of Eigenvalues::calculate_gradients (src/Eigenvalues.cpp:46-70): the scaled Rayleigh quotient
        of the Hessian at rho (plus vshift (vshift.x)/2) and its gradient; ideal = full_hessian_."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        if rho.size != self.N or x.size != self.N:
            raise ValueError(f"expected {self.N} densities and components")
        vs = None if vshift is None else np.ascontiguousarray(vshift, dtype=np.float64)
        df = np.zeros(x.shape)
        f = _d()
        fl = (0 if ideal else NO_IDEAL) | (SQUARED_MATRIX if squared else 0) | (DYNAMIC_MATRIX if dynamic else 0)
        _check(lib().cdft_system_eigen_objective(self.h, _ptr(rho), _ptr(x), None if vs is None else _ptr(vs),
                                                 float(scale), int(bool(change_sign)), fl, ctypes.byref(f),
                                                 _ptr(df)))
        return f.value, df

    def log_det(self, rho, order, lam_max, lam_min, seed, num_samples, has_zero_eigenvalue=False, ideal=True,
                squared=False, dynamic=False):
        """(log det, variance): Log_Det(dft, order, lam_max, lam_min).calculate_log_det(seed, num_samples,
        has_zero_eigenvalue, variance) (src/Log_Det.cpp:22-166); ideal = full_hessian_, squared =
        use_squared_matrix_."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.size != self.N:
            raise ValueError(f"expected {self.N} densities")
        ld, var = _d(), _d()
        fl = (0 if ideal else NO_IDEAL) | (SQUARED_MATRIX if squared else 0) | (DYNAMIC_MATRIX if dynamic else 0)
        _check(lib().cdft_system_log_det(self.h, _ptr(rho), int(order), float(lam_max), float(lam_min), int(seed),
                                         int(num_samples), int(bool(has_zero_eigenvalue)), fl, ctypes.byref(ld),
                                         ctypes.byref(var)))
        return ld.value, var.value

    def fmt_hessian_dot(self, rho, v, out=None, accumulate=False):
        """FMT::add_second_derivative: the hard-sphere part of the Hessian applied to v."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if rho.size != self.nb * self.N or v.size != rho.size:
            raise ValueError(f"expected {self.nb * self.N} densities and directions")
        res = _out(out, self.nb * self.N, self._ashape())
        _check(lib().cdft_fmt_hessian_dot(self.h, _ptr(rho), _ptr(v), _ptr(res), ACCUMULATE if accumulate else 0))
        return res

    def autotune(self, reps=3):
        """cdft_system_autotune: picks the fastest kernel tiles for this shape; returns ms per evaluation."""
        ms = _d()
        _check(lib().cdft_system_autotune(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    # -- evaluations on device pointers (ints) --
    def eval_device(self, rho_ptr, dF_ptr, parts=None, fmt_only=False, accumulate=False):
        flags = DEVICE_PTRS | (ACCUMULATE if accumulate else 0)
        if fmt_only:
            F = _d()
            _check(lib().cdft_fmt_eval(self.h, _p(rho_ptr), _p(dF_ptr), ctypes.byref(F), flags))
            return F.value
        p = np.zeros(5) if parts is None else parts
        _check(lib().cdft_system_eval(self.h, _p(rho_ptr), _p(dF_ptr), _ptr(p), flags))
        return p

    def eval_async(self, rho_ptr, dF_ptr, fmt_only=False, accumulate=False):
        flags = (FMT_ONLY if fmt_only else 0) | (ACCUMULATE if accumulate else 0)
        _check(lib().cdft_system_eval_async(self.h, _p(rho_ptr), _p(dF_ptr), flags))

    def fetch(self):
        p = np.zeros(5 * self.nb)
        _check(lib().cdft_system_fetch(self.h, _ptr(p)))
        return p if self.nb == 1 else p.reshape(self.nb, 5)

    def time(self, rho_ptr, dF_ptr, iters, fmt_only=False):
        """Milliseconds for `iters` back-to-back evaluations (hipEvents on the system stream)."""
        ms = _d()
        flags = FMT_ONLY if fmt_only else 0
        _check(lib().cdft_system_time(self.h, _p(rho_ptr), _p(dF_ptr), flags, int(iters), ctypes.byref(ms)))
        return ms.value

    @property
    def stream(self):
        s = _p()
        _check(lib().cdft_system_stream(self.h, ctypes.byref(s)))
        return s.value or 0

    def set_stream(self, hip_stream):
        """cdft_system_set_stream: later work runs on hip_stream (an address; 0 / None: the system's own)."""
        _check(lib().cdft_system_set_stream(self.h, _p(hip_stream or None)))

    def chem_potential(self):
        """Species::getChemPotential after the last evaluation: the fixed-mass multiplier (or mu); (nb,) for
        batches."""
        mu = np.zeros(self.nb)
        _check(lib().cdft_system_chem_potential(self.h, _ptr(mu)))
        return mu[0] if self.nb == 1 else mu

    def bulk(self, rho):
        """(beta f, beta mu) at uniform density: DFT::fhelmholtz_times_beta_over_volume, mu_times_beta."""
        f, mu = _d(), _d()
        _check(lib().cdft_system_bulk(self.h, float(rho), ctypes.byref(f), ctypes.byref(mu)))
        return f.value, mu.value

    @property
    def a_vdw(self):
        a = _d()
        _check(lib().cdft_system_a_vdw(self.h, ctypes.byref(a)))
        return a.value

    @property
    def layout(self):
        """(stored weighted-density arrays, convolved arrays) of the pipeline (cdft_system_layout)."""
        a, b = _i(), _i()
        _check(lib().cdft_system_layout(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def device_weights(self):
        w = np.zeros((11,) + self.shape)
        _check(lib().cdft_system_get_weights(self.h, _ptr(w)))
        return w

    def device_mf_kernel(self):
        w = np.zeros(self.shape)
        _check(lib().cdft_system_get_mf_kernel(self.h, _ptr(w)))
        return w

    def set_timing(self, on=True):
        _check(lib().cdft_system_set_timing(self.h, int(bool(on))))

    def timing(self):
        n = _i(0)
        ms = np.zeros(16)
        names = (ctypes.c_char_p * 16)()
        _check(lib().cdft_system_get_timing(self.h, ctypes.byref(n), _ptr(ms), names))
        return [(names[i].decode(), float(ms[i])) for i in range(n.value)]


class Fire:
    """fireMinimizer2 on the device (src/Minimizer.cpp:117-327)."""

    def __init__(self, system, rho0, dt=None, dt_max=None, **params):
        self.system = system
        p = FireParams()
        _check(lib().cdft_fire_default_params(ctypes.byref(p)))
        if dt is not None:
            p.dt = dt
            p.dt_max = 10 * dt if dt_max is None else dt_max
        elif dt_max is not None:
            p.dt_max = dt_max
        for k, v in params.items():
            setattr(p, k, v)
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64)
        h = _p()
        if isinstance(system, Mixture):  # one FIRE2 state over all species (rho0: (nspecies,) + shape)
            S = system.nspecies
            if rho0.size != S * system.N:
                raise ValueError(f"expected {S} x {system.N} densities")
            rho0 = rho0.reshape((S,) + system.shape)
            rp = (_p * S)(*[rho0[i].ctypes.data for i in range(S)])
            _check(lib().cdft_mixture_fire_create(system.h, ctypes.cast(rp, _p), 0, ctypes.byref(p),
                                                  ctypes.byref(h)))
        else:
            _check(lib().cdft_fire_create(system.h, _ptr(rho0), 0, ctypes.byref(p), ctypes.byref(h)))
        self.h = h

    def step(self):
        """Returns (status, report): status 0 ok, 1 back-tracked (eta > 1), 2 uphill limit."""
        r = FireReport()
        rc = lib().cdft_fire_step(self.h, ctypes.byref(r))
        if rc < 0:
            _check(rc)
        return rc, r.as_dict()

    def run(self, max_steps, tol_raw):
        r = FireReport()
        rc = lib().cdft_fire_run(self.h, int(max_steps), float(tol_raw), ctypes.byref(r))
        if rc < 0:
            _check(rc)
        return rc, r.as_dict()

    def state(self):
        shp = self.system.shape
        if isinstance(self.system, Mixture):
            shp = (self.system.nspecies,) + shp
        rho, df, x, v = (np.zeros(shp) for _ in range(4))
        _check(lib().cdft_fire_get(self.h, _ptr(rho), _ptr(df), _ptr(x), _ptr(v)))
        return rho, x, v, df

    def set_fixed_direction(self, e):
        """Minimizer::set_fixed_direction: each force is reflected, dF -= 2 (e . dF) e (None clears)."""
        d = None if e is None else np.ascontiguousarray(e, dtype=np.float64)
        _check(lib().cdft_fire_set_fixed_direction(self.h, _ptr(d), 0))

    def reset(self, rho):
        """fireMinimizer2::reset from the density rho (dt, alpha and counters kept)."""
        r = np.ascontiguousarray(rho, dtype=np.float64)
        return lib().cdft_fire_reset(self.h, _ptr(r), 0)

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_fire_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DDFT:
    """Integrating-factor DDFT on the device (src/DDFT.cpp:42-351)."""

    def __init__(self, system, rho0, dt=0.0, dt_max=0.0, tol=0.0, central_differences=False):
        """central_differences: the third argument of DDFT(dft, showGraphics, central_differences)."""
        self.system = system
        rho0 = np.ascontiguousarray(rho0, dtype=np.float64)
        h = _p()
        _check(lib().cdft_ddft_create(system.h, _ptr(rho0), float(dt), float(dt_max), float(tol), ctypes.byref(h)))
        self.h = h
        if central_differences:
            _check(lib().cdft_ddft_set_central_differences(self.h, 1))

    def step(self):
        r = DdftReport()
        _check(lib().cdft_ddft_step(self.h, ctypes.byref(r)))
        return r.as_dict()

    def set_max_time_step(self, dt_max):
        _check(lib().cdft_ddft_set_max_time_step(self.h, float(dt_max)))

    def density(self):
        out = np.zeros(self.system.shape)
        _check(lib().cdft_ddft_get_density(self.h, _ptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_ddft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FireBatch:
    """FIRE2 on nb same-shape systems with device-side control (cdft_fire_batch_*)."""

    def __init__(self, shape, dx, mus, rho0, vexts=None, hsd=1.0, model=WHITEBEAR_I, precision=FP64, ideal=True,
                 mean_field=False, sigma=1.0, eps=1.0, rcut=2.5, kT=1.0, A=1.0, B=0.0, dt=None, dt_max=None,
                 **params):
        self.shape = tuple(int(n) for n in shape)
        dxs = (dx, dx, dx) if np.isscalar(dx) else tuple(dx)
        self.N = int(np.prod(self.shape))
        m = np.ascontiguousarray(mus, dtype=np.float64).reshape(-1)
        self.nb = m.size
        r0 = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        if r0.size != self.nb * self.N:
            raise ValueError("rho0 must hold nb x N values")
        vx = None if vexts is None else np.ascontiguousarray(vexts, dtype=np.float64).reshape(-1)
        d = _Desc(self.shape[0], self.shape[1], self.shape[2], dxs[0], dxs[1], dxs[2], int(model), float(A), float(B),
                  float(hsd), int(precision), None, int(bool(ideal)), 0.0, None, int(bool(mean_field)), float(sigma),
                  float(eps), float(rcut), float(kT), None, 0, 0.0, 0, 0.0, 0, 0, 0.0, 0.0, None, 0.0, 0, 0.0)
        p = FireParams()
        _check(lib().cdft_fire_default_params(ctypes.byref(p)))
        if dt is not None:
            p.dt = dt
            p.dt_max = 10 * dt if dt_max is None else dt_max
        elif dt_max is not None:
            p.dt_max = dt_max
        for k, v in params.items():
            setattr(p, k, v)
        h = _p()
        _check(lib().cdft_fire_batch_create(ctypes.byref(d), self.nb, m.ctypes.data_as(_p), _ptr(vx), _ptr(r0),
                                            ctypes.byref(p), ctypes.byref(h)))
        self.h = h

    def run(self, steps, tol_raw, graph=True):
        """Enqueue `steps` steps (no synchronisation)."""
        _check(lib().cdft_fire_batch_run(self.h, int(steps), float(tol_raw), int(bool(graph))))

    def autotune(self, reps=3):
        """cdft_system_autotune on the batch's current system; ms per evaluation of all members."""
        sysh = _p()
        _check(lib().cdft_fire_batch_system(self.h, ctypes.byref(sysh)))
        ms = _d()
        _check(lib().cdft_system_autotune(sysh, int(reps), ctypes.byref(ms)))
        return ms.value

    def set_precision(self, precision):
        """Evaluate with FP64 or MIXED from now on (re-evaluates the current densities)."""
        _check(lib().cdft_fire_batch_set_precision(self.h, int(precision)))

    def status(self):
        """(reports, status array, number still running); synchronises."""
        reps = (FireReport * self.nb)()
        st = np.zeros(self.nb, dtype=np.int32)
        run = _i()
        _check(lib().cdft_fire_batch_status(self.h, ctypes.cast(reps, _p), st.ctypes.data_as(_p), ctypes.byref(run)))
        return [r.as_dict() for r in reps], st, run.value

    def density(self, member):
        out = np.zeros(self.shape)
        _check(lib().cdft_fire_batch_get(self.h, int(member), _ptr(out)))
        return out

    def state(self, member):
        """(rho, x, v, dF_raw) of one member, as Fire.state (cdft_fire_batch_state)."""
        rho, df, x, v = (np.zeros(self.shape) for _ in range(4))
        _check(lib().cdft_fire_batch_state(self.h, int(member), _ptr(rho), _ptr(df), _ptr(x), _ptr(v)))
        return rho, x, v, df

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_fire_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FireGroup:
    """Several FireBatch objects (e.g. one per grid size) stepped together: one hipGraph launch per `run` advances
    every batch, each on its own stream inside the graph (cdft_fire_group_*).  The batches stay usable on their own
    (status, density, set_precision); close the group before them."""

    def __init__(self, batches):
        self.batches = list(batches)
        arr = (_p * len(self.batches))(*[b.h for b in self.batches])
        h = _p()
        _check(lib().cdft_fire_group_create(arr, len(self.batches), ctypes.byref(h)))
        self.h = h

    def run(self, steps, tol_raw):
        """Enqueue `steps` steps of every batch as one graph launch (no synchronisation)."""
        _check(lib().cdft_fire_group_run(self.h, int(steps), float(tol_raw)))

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_fire_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mixture:
    """Hard spheres of several diameters on one lattice (cdft_mixture_*): the FMT of a mixture,
    FMT::calculateFreeEnergyAndDerivatives over several FMT species (src/FMT.cpp:363-427).  FP64."""

    def __init__(self, shape, dx, hsd, model=WHITEBEAR_I, weights=None, A=1.0, B=0.0,
                 weights_mode=WEIGHTS_ACCURATE, ideal=True, mu=None, vext=None, fixed_mass=None, boundary=None):
        """hsd: one diameter per species; weights: None, or one entry per species (None or 11*N cell weights).
        ideal, mu (one beta*mu per species), vext (None, or one entry per species: None or N values of
        beta*V_ext), fixed_mass (None or one value per species, 0: free) and boundary (None or one flag set per
        species) are the rest of the DFT that eval and FIRE over the mixture use."""
        self.shape = tuple(int(n) for n in shape)
        dxs = (dx, dx, dx) if np.isscalar(dx) else tuple(dx)
        self.dx = dxs
        self.N = int(np.prod(self.shape))
        self._hsd = np.ascontiguousarray(np.atleast_1d(hsd), dtype=np.float64).reshape(-1)
        self.nspecies = int(self._hsd.size)
        if self.nspecies < 1:
            raise ValueError("a mixture needs at least one species")
        wptr = None
        self._w = []
        if weights is not None:
            if len(weights) != self.nspecies:
                raise ValueError(f"expected {self.nspecies} weight sets")
            ptrs = (_p * self.nspecies)()
            for i, w in enumerate(weights):
                if w is None:
                    ptrs[i] = None
                    continue
                a = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
                if a.size != 11 * self.N:
                    raise ValueError(f"expected {11 * self.N} cell weights per species")
                self._w.append(a)
                ptrs[i] = a.ctypes.data
            self._wptrs = ptrs
            wptr = ctypes.cast(ptrs, _p)
        mptr = None
        if mu is not None:
            self._mu = np.ascontiguousarray(np.atleast_1d(mu), dtype=np.float64).reshape(-1)
            if self._mu.size != self.nspecies:
                raise ValueError(f"expected {self.nspecies} chemical potentials")
            mptr = self._mu.ctypes.data
        vptr = None
        self._v = []
        if vext is not None:
            if len(vext) != self.nspecies:
                raise ValueError(f"expected {self.nspecies} external fields")
            vp = (_p * self.nspecies)()
            for i, v in enumerate(vext):
                if v is None:
                    vp[i] = None
                    continue
                a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                if a.size != self.N:
                    raise ValueError(f"expected {self.N} external-field values per species")
                self._v.append(a)
                vp[i] = a.ctypes.data
            self._vptrs = vp
            vptr = ctypes.cast(vp, _p)
        fptr = bptr = None
        if fixed_mass is not None:
            self._fm = np.ascontiguousarray(np.atleast_1d(fixed_mass), dtype=np.float64).reshape(-1)
            if self._fm.size != self.nspecies:
                raise ValueError(f"expected {self.nspecies} fixed masses")
            fptr = self._fm.ctypes.data
        if boundary is not None:
            self._bnd = np.ascontiguousarray(np.atleast_1d(boundary), dtype=np.int32).reshape(-1)
            if self._bnd.size != self.nspecies:
                raise ValueError(f"expected {self.nspecies} boundary flag sets")
            bptr = self._bnd.ctypes.data
        d = _MixDesc(self.shape[0], self.shape[1], self.shape[2], dxs[0], dxs[1], dxs[2], int(model), float(A),
                     float(B), self.nspecies, self._hsd.ctypes.data, wptr, int(weights_mode), int(bool(ideal)), mptr,
                     vptr, fptr, bptr)
        h = _p()
        _check(lib().cdft_mixture_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self._w = []
        self._wptrs = None
        self._v = []
        self._vptrs = None
        self.model = model

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_mixture_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fmt_eval(self, rho, dF=None, accumulate=False):
        """Returns (F_hs, dF) for rho of shape (nspecies,) + shape; dF has rho's shape (written, or added to)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.size != self.nspecies * self.N:
            raise ValueError(f"expected {self.nspecies} x {self.N} densities, got {rho.size}")
        rho = rho.reshape((self.nspecies,) + self.shape)
        out = _out(dF, self.nspecies * self.N, (self.nspecies,) + self.shape)
        out = out.reshape((self.nspecies,) + self.shape)
        rp = (_p * self.nspecies)(*[rho[i].ctypes.data for i in range(self.nspecies)])
        fp = (_p * self.nspecies)(*[out[i].ctypes.data for i in range(self.nspecies)])
        F = _d(0.0)
        _check(lib().cdft_mixture_fmt_eval(self.h, ctypes.cast(rp, _p), ctypes.cast(fp, _p), ctypes.byref(F),
                                           ACCUMULATE if accumulate else 0))
        return F.value, out

    def eval(self, rho, dF=None, accumulate=False, only_fex=False):
        """DFT::calculateFreeEnergyAndDerivatives over the species: returns (parts[5], dF), parts = {F, F_id, F_hs,
        0, F_ext}; only_fex leaves out the ideal-gas terms."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.size != self.nspecies * self.N:
            raise ValueError(f"expected {self.nspecies} x {self.N} densities, got {rho.size}")
        rho = rho.reshape((self.nspecies,) + self.shape)
        out = _out(dF, self.nspecies * self.N, (self.nspecies,) + self.shape).reshape((self.nspecies,) + self.shape)
        rp = (_p * self.nspecies)(*[rho[i].ctypes.data for i in range(self.nspecies)])
        fp = (_p * self.nspecies)(*[out[i].ctypes.data for i in range(self.nspecies)])
        parts = np.zeros(5)
        flags = (ACCUMULATE if accumulate else 0) | (NO_IDEAL if only_fex else 0)
        _check(lib().cdft_mixture_eval(self.h, ctypes.cast(rp, _p), ctypes.cast(fp, _p), _ptr(parts), flags))
        return parts, out

    def chem_potential(self):
        """cdft_mixture_chem_potential: each species' fixed-mass multiplier after the last evaluation, or its mu."""
        mu = np.zeros(self.nspecies)
        _check(lib().cdft_mixture_chem_potential(self.h, _ptr(mu)))
        return mu

    def fmt_eval_device(self, rho_ptrs, dF_ptrs, accumulate=False):
        """Device pointers (one per species): returns F_hs; dF written or added to on the device."""
        if len(rho_ptrs) != self.nspecies or len(dF_ptrs) != self.nspecies:
            raise ValueError(f"expected {self.nspecies} density and force pointers")
        rp = (_p * self.nspecies)(*rho_ptrs)
        fp = (_p * self.nspecies)(*dF_ptrs)
        F = _d(0.0)
        _check(lib().cdft_mixture_fmt_eval(self.h, ctypes.cast(rp, _p), ctypes.cast(fp, _p), ctypes.byref(F),
                                           DEVICE_PTRS | (ACCUMULATE if accumulate else 0)))
        return F.value


class DeviceArray:
    """A device buffer (float64 by default) allocated through the C ABI (no second HIP runtime)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _p()
        _check(lib().cdft_malloc(ctypes.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        d.upload(a)
        return d

    def upload(self, a, offset=0):
        """Copy host array a into the buffer, `offset` elements in."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert offset * self.dtype.itemsize + a.nbytes <= self.nbytes
        _check(lib().cdft_memcpy_h2d(_p(self.ptr + offset * self.dtype.itemsize), _ptr(a), a.nbytes))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().cdft_memcpy_d2h(_ptr(out), _p(self.ptr), self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().cdft_free(_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class UNet:
    """The warm-start network (cdft_unet_*): (beta V_ext, mu) on D^3 grids -> log rho."""

    def __init__(self, D, seed=1):
        h = _p()
        _check(lib().cdft_unet_create(int(D), ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self.h = h
        self.D = int(D)
        n = ctypes.c_longlong()
        _check(lib().cdft_unet_num_params(h, ctypes.byref(n)))
        self.num_params = n.value

    def params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        _check(lib().cdft_unet_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.num_params
        _check(lib().cdft_unet_set_params(self.h, _ptr(w)))

    def _idx(self, idx, *arrays):
        """Sample numbers as int32, checked against the sample count of every data array."""
        i = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1)
        if i.size == 0:
            raise ValueError("empty batch")
        for a in arrays:
            n = a.shape[0]
            if a.dtype != np.float32 or i.min() < 0 or i.max() >= n:
                raise ValueError("sample index out of range or data not float32")
        for a in arrays[:1] + arrays[2:]:
            if a.shape[1:] != (self.D,) * 3:
                raise ValueError("grids must be D^3")
        return i

    def forward(self, V, mu, idx, out):
        """V, mu, out: DeviceArray (float32); idx: sample numbers of the batch."""
        i = self._idx(idx, V, mu)
        if out.shape != (i.size,) + (self.D,) * 3 or out.dtype != np.float32:
            raise ValueError("out must be float32 (nb, D, D, D)")
        _check(lib().cdft_unet_forward(self.h, _p(V.ptr), _p(mu.ptr), _ptr(i), i.size, _p(out.ptr)))

    def grad(self, V, mu, T, idx):
        i = self._idx(idx, V, mu, T)
        loss = ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32)
        _check(lib().cdft_unet_grad(self.h, _p(V.ptr), _p(mu.ptr), _p(T.ptr), _ptr(i), i.size, ctypes.byref(loss),
                                    _ptr(g)))
        return loss.value, g

    def train_step(self, V, mu, T, idx, lr):
        i = self._idx(idx, V, mu, T)
        loss = ctypes.c_float()
        _check(lib().cdft_unet_train_step(self.h, _p(V.ptr), _p(mu.ptr), _p(T.ptr), _ptr(i), i.size, float(lr),
                                          ctypes.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _pot(kind, sigma, eps, rcut, alpha, bh, r_att_min):
    return _PotDesc(int(kind), float(sigma), float(eps), float(rcut), float(alpha), int(bool(bh)), float(r_att_min))


def potential_hsd(kT, kind=POT_LJ, sigma=1.0, eps=1.0, rcut=2.5, alpha=0.0, bh=False, r_att_min=0.0):
    """Potential1::getHSD for any potential and split (cdft_potential_hsd)."""
    out = _d()
    _check(lib().cdft_potential_hsd(ctypes.byref(_pot(kind, sigma, eps, rcut, alpha, bh, r_att_min)), float(kT),
                                    ctypes.byref(out)))
    return out.value


def potential_vdw(kT, kind=POT_LJ, sigma=1.0, eps=1.0, rcut=2.5, alpha=0.0, bh=False, r_att_min=0.0):
    """Potential1::getVDW_Parameter (cdft_potential_vdw)."""
    out = _d()
    _check(lib().cdft_potential_vdw(ctypes.byref(_pot(kind, sigma, eps, rcut, alpha, bh, r_att_min)), float(kT),
                                    ctypes.byref(out)))
    return out.value


def potential_watt(r, kind=POT_LJ, sigma=1.0, eps=1.0, rcut=2.5, alpha=0.0, bh=False, r_att_min=0.0):
    """Potential1::Watt at the distances r (cdft_potential_watt)."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    w = np.empty_like(r)
    _check(lib().cdft_potential_watt(ctypes.byref(_pot(kind, sigma, eps, rcut, alpha, bh, r_att_min)), r.size, _ptr(r),
                                     _ptr(w)))
    return w


def lj_hsd(kT, sigma=1.0, eps=1.0, rcut=2.5):
    """Hard-sphere diameter of the WCA-split LJ potential (Potential1::getHSD, include/Potential1.h:73-83)."""
    d = _d()
    _check(lib().cdft_lj_hsd(float(sigma), float(eps), float(rcut), float(kT), ctypes.byref(d)))
    return d.value


def fft_r2c(x, precision=FP64):
    x = np.ascontiguousarray(x, dtype=np.float64)
    Nx, Ny, Nz = x.shape
    out = np.zeros((Nx, Ny, Nz // 2 + 1), dtype=np.complex128)
    _check(lib().cdft_fft_r2c(Nx, Ny, Nz, precision, _ptr(x), _ptr(out)))
    return out


def fft_c2r(s, Nz, precision=FP64):
    s = np.ascontiguousarray(s, dtype=np.complex128)
    Nx, Ny, _ = s.shape
    out = np.zeros((Nx, Ny, Nz))
    _check(lib().cdft_fft_c2r(Nx, Ny, Nz, precision, _ptr(s), _ptr(out)))
    return out


class OpenDDFT:
    """Open-boundary DDFT on the device (DDFT_IF_Open, mduran/src/DDFT_IF_Open.cpp:25-463): the density is held
    at `background` on the planes ix, iy or iz = 0, the interior evolves through 3D DST-I transforms."""

    def __init__(self, system, rho0, background, dt=0.0, dt_max=0.0, tol=0.0):
        self.system = system
        r = np.ascontiguousarray(rho0, dtype=np.float64)
        h = _p()
        _check(lib().cdft_ddft_open_create(system.h, _ptr(r), float(background), float(dt), float(dt_max), float(tol),
                                           ctypes.byref(h)))
        self.h = h

    def step(self):
        r = OpenDdftReport()
        _check(lib().cdft_ddft_open_step(self.h, ctypes.byref(r)))
        return r.as_dict()

    def density(self):
        rho = np.zeros(self.system.shape)
        _check(lib().cdft_ddft_open_get_density(self.h, _ptr(rho)))
        return rho

    def close(self):
        if getattr(self, "h", None):
            lib().cdft_ddft_open_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dst1_3d(a):
    """3D DST-I (FFTW RODFT00) of a compact array on the GPU."""
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    _check(lib().cdft_dst1_3d(*a.shape, _ptr(a)))
    return a


def mf_kernel_points(shape, dx, points, potential=POT_LJ, sigma=1.0, eps=1.0, rcut=2.5, alpha=0.0, bh=False,
                     r_att_min=0.0, nradial=0):
    """Points-file mean-field kernel on the GPU (Interaction::generateWeights, src/Interaction.cpp:672-860), before
    the division by kT: pass kernel / kT as System(mf_kernel=...).  points: (n, 4) rows (ux, uy, uz, weight) as
    formats.read_points_file returns them."""
    dxs = (dx, dx, dx) if np.isscalar(dx) else tuple(dx)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 4)
    pd = _PotDesc(int(potential), float(sigma), float(eps), float(rcut), float(alpha), int(bool(bh)), float(r_att_min))
    out = np.zeros(tuple(int(n) for n in shape))
    _check(lib().cdft_mf_kernel_points(int(shape[0]), int(shape[1]), int(shape[2]), dxs[0], dxs[1], dxs[2],
                                       ctypes.byref(pd), _ptr(pts), len(pts), int(nradial), _ptr(out)))
    return out


class Eigenvalues:
    """Eigenvalues(matrix) of include/Eigenvalues.h:16-59: the smallest (or, with change_sign, the largest)
    eigenvalue of the Hessian of `system` at `rho` by minimising the scaled Rayleigh quotient of
    System.eigen_objective (every product on the GPU).  The reference minimises with nlopt's LD_LBFGS, absent here;
    this uses scipy's L-BFGS-B with the same relative tolerance on f, so the iterates differ and the converged
    eigenpair agrees."""

    def __init__(self, system, rho, ideal=True, squared=False, verbose=False, dynamic=False):
        self.system, self.rho = system, np.ascontiguousarray(rho, dtype=np.float64)
        self.ideal, self.squared, self.verbose, self.dynamic = ideal, squared, verbose, dynamic
        self.scale, self.tol, self.change_sign, self.vshift = 1e6, 1e-6, False, None
        self.eigen_vec, self.eigen_val, self.num_eval, self.return_code = None, 0.0, 0, 0

    def set_scale(self, d):
        self.scale = float(d)

    def set_tolerence(self, d):
        self.tol = float(d)

    def set_change_sign(self, b):
        self.change_sign = bool(b)

    def set_vshift(self, v):
        self.vshift = np.ascontiguousarray(v, dtype=np.float64)

    def calculate_gradients(self, x):
        """(f, df) at x (src/Eigenvalues.cpp:46-70)."""
        self.num_eval += 1
        return self.system.eigen_objective(self.rho, x, self.vshift, self.scale, self.change_sign, self.ideal,
                                           self.squared, self.dynamic)

    def calculate_eigenvector(self, x0=None, seed=0, maxiter=10000):
        """src/Eigenvalues.cpp:72-99: start from x0 (else a seeded normal vector, zero on a fixed boundary)."""
        from scipy.optimize import minimize
        if x0 is None:
            x0 = np.random.default_rng(seed).standard_normal(self.rho.size)
            if self.system.fixed_background:
                m = np.zeros(self.system.shape, bool)
                m[0], m[:, 0], m[:, :, 0] = True, True, True
                x0[m.ravel()] = 0.0
        res = minimize(lambda x: self.calculate_gradients(x), np.ravel(x0).astype(np.float64), jac=True,
                       method="L-BFGS-B", options={"ftol": self.tol, "gtol": 0.0, "maxiter": maxiter,
                                                   "maxfun": maxiter})
        self.return_code = int(res.status)
        self.eigen_val = float(res.fun) / self.scale * (-1.0 if self.change_sign else 1.0)
        self.eigen_vec = res.x / np.linalg.norm(res.x)
        return self.eigen_val


class Arnoldi:
    """Arnoldi(matrix) of include/Arnoldi.h:18-84: the k eigenpairs of A + shift I of largest modulus, reported as
    eigenvalue real part minus shift (src/Arnoldi.cpp:205-211, ordered as compute_and_sort_eigenvectors :286-306),
    with A the Hessian of `system` at `rho` or, dynamic=True, the DDFT dynamical matrix (System.dynamic_dot).  Every
    product runs on the GPU; complex vectors are applied as two real products and boundary entries are zeroed with a
    fixed background, as Arnoldi::matrix_dot_v does (src/Arnoldi.cpp:25-70).  The reference restarts its own
    factorisation with armadillo (absent here); this drives ARPACK through scipy with k + p basis vectors, so the
    iterates differ and the converged eigenpairs agree."""

    def __init__(self, system, rho, ideal=True, dynamic=False, verbose=False):
        self.system, self.rho = system, np.ascontiguousarray(rho, dtype=np.float64)
        self.ideal, self.dynamic, self.verbose = ideal, dynamic, verbose
        self.eigenvalues, self.eigenvectors, self.success, self.executed = [], [], False, False
        self.initial_guess = None
        n = self.rho.size
        self.mask = np.zeros(system.shape, bool)
        if system.fixed_background:
            self.mask[0], self.mask[:, 0], self.mask[:, :, 0] = True, True, True
        self.mask = self.mask.ravel()
        self.n = n

    def set_initial_guess(self, v):
        v = np.ravel(np.asarray(v, dtype=np.float64))
        self.initial_guess = v / np.linalg.norm(v)

    def _real_dot(self, x):
        x = np.where(self.mask, 0.0, x).reshape(self.system.shape)
        if self.dynamic:
            return self.system.dynamic_dot(self.rho, x, ideal=self.ideal).ravel()
        return self.system.hessian_dot(self.rho, x, ideal=self.ideal).ravel()

    def matrix_dot_v(self, v, shift):
        """(A + shift I) v for a complex v (src/Arnoldi.cpp:25-70)."""
        v = np.asarray(v)
        out = self._real_dot(v.real).astype(np.complex128)
        if np.iscomplexobj(v) and np.any(v.imag):
            out += 1j * self._real_dot(v.imag)
        out = out + shift * v
        out[self.mask] = 0.0
        return out

    def determine_largest_eigenpairs(self, shift, k, p, compute_all_k_largest=True, maxSteps=1000000, tol=1e-8):
        from scipy.sparse.linalg import LinearOperator, eigs
        op = LinearOperator((self.n, self.n), matvec=lambda v: self.matrix_dot_v(v, shift), dtype=np.complex128)
        ncv = min(self.n - 1, k + p)
        w, V = eigs(op, k=k, which="LM", ncv=ncv, tol=tol, maxiter=int(min(maxSteps, 2 ** 31 - 1)),
                    v0=None if self.initial_guess is None else self.initial_guess.astype(np.complex128))
        order = np.argsort(-np.abs(w), kind="stable")
        self.eigenvalues = [float(w[i].real - shift) for i in order]
        self.eigenvectors = [V[:, i] for i in order]
        self.executed, self.success = True, True
        return self.eigenvalues

    def is_success(self):
        if not self.executed:
            raise RuntimeError("Arnoldi: Solver has yet not been executed")
        return self.success

    def get_eigenvalue(self, i):
        if not self.executed:
            raise RuntimeError("Arnoldi: Solver has yet not been executed")
        return self.eigenvalues[i]

    def get_eigenvector(self, i):
        if not self.executed:
            raise RuntimeError("Arnoldi: Solver has yet not been executed")
        return self.eigenvectors[i]
