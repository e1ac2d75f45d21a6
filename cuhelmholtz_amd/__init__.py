"""cuhelmholtz_amd -- MI355X (gfx950) 3D mixed DFT/DCT/DST transforms and
FFT-diagonalised Helmholtz solves.

Thin Python mirror of the C ABI (include/cuhelmholtz.h) over ctypes.  The
arrays are torch tensors on the GPU (torch is only the device-memory and
stream plumbing); every call goes to the HIP library built in-tree
(libcuhelmholtz_amd.so).  There is no CPU fallback: if the library is missing
the import of a compute entry point raises.

Reference API mirrored (names, argument meaning, conventions):
  do_transform(d_data, d_out, NX, NY, NZ, X, Y, Z)  cudasymmfft/funcinterface.h:5-12
  transform_kind ordinals                            cudasymmfft/global.h:6-15
  cuda_helmholtz_solver(...)                         cudahelmholtz/cuda_helmholtz_solver.h:5-11
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuhelmholtz_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

DCT_1, DCT_2, DCT_3, DST_1, DST_2, DST_3, DFT_R2C, DFT_C2R = range(8)
KIND_NAMES = ["DCT_1", "DCT_2", "DCT_3", "DST_1", "DST_2", "DST_3", "DFT_R2C", "DFT_C2R"]

# bc code -> (forward kind, inverse kind), cudahelmholtz/cuda_helmholtz_solver.cu:78-172
BC_KINDS = {0: (DFT_R2C, DFT_C2R), 1: (DST_1, DST_1), 2: (DST_3, DST_2), 3: (DCT_1, DCT_1), 4: (DCT_3, DCT_2)}


class CHZError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CHZError(f"{LIB_PATH} is missing: run `python -m cuhelmholtz_amd.build` (needs hipcc)")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, fp, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_float, ctypes.c_int
        L.chz_do_transform.argtypes = [vp, vp, i, i, i, i, i, i, vp]
        L.chz_do_transform_f32.argtypes = [vp, vp, i, i, i, i, i, i, vp]
        L.chz_last_error.restype = ctypes.c_char_p
        L.chz_set_timing.argtypes = [i]
        L.chz_set_axis_config.argtypes = [i] * 6
        L.chz_load_tuning.argtypes = [ctypes.c_char_p]
        L.chz_radix_plan.argtypes = [i, i, i, ctypes.POINTER(ctypes.c_ulonglong)]
        L.chz_compiled_kernels.argtypes = [ctypes.POINTER(i), i]
        L.chz_get_timing.argtypes = [ctypes.POINTER(dp), i, ctypes.c_char_p, i]
        L.chz_axis_config.argtypes = [i, i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                                      ctypes.POINTER(ctypes.c_ulonglong)]
        _lib = L
        _declare_solver(L)
        _declare_batch(L)
    return _lib


def _declare_solver(L):
    vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    if hasattr(L, "chz_helmholtz_solve"):
        L.chz_helmholtz_solve.argtypes = ([d, d, i, i, vp, vp] * 3) + [d, vp, vp]
    if hasattr(L, "chz_helmholtz_solve19"):
        L.chz_helmholtz_solve19.argtypes = ([d, d, i, i, vp, vp] * 3) + [d, vp, vp]
    if hasattr(L, "chz_helmholtz_solve_noalloc"):
        L.chz_helmholtz_solve_noalloc.argtypes = ([d, d, i, i, vp, vp] * 3) + [d, vp, vp]
    if hasattr(L, "chz_helmholtz_solve_mixed"):
        L.chz_helmholtz_solve_mixed.argtypes = ([d, d, i, i, vp, vp] * 3) + [d, vp, i, vp]
    if hasattr(L, "chz_helmholtz_adjoint"):
        L.chz_helmholtz_adjoint.argtypes = ([d, d, i, i] * 3) + [d, vp, vp, vp] + [vp] * 6 + [vp, vp]
    if hasattr(L, "chz_helmholtz_adjoint_batch"):
        L.chz_helmholtz_adjoint_batch.argtypes = ([d, d, i, i] * 3) + [i, vp, vp, vp, vp] + [vp] * 6 + [vp, vp]


class _Problem(ctypes.Structure):
    _fields_ = [("domain", ctypes.c_double * 6), ("N", ctypes.c_int * 3), ("bc", ctypes.c_int * 3),
                ("lam", ctypes.c_double), ("f", ctypes.c_void_p), ("faces", ctypes.c_void_p * 6)]


def _declare_batch(L):
    vp, i = ctypes.c_void_p, ctypes.c_int
    L.chz_batch_create.argtypes = [ctypes.POINTER(_Problem), i, i, i, ctypes.POINTER(vp)]
    L.chz_batch_solve.argtypes = [vp, i, vp]
    L.chz_batch_rebind.argtypes = [vp, ctypes.POINTER(_Problem), i, vp]
    L.chz_batch_solve_cached.argtypes = [ctypes.POINTER(_Problem), i, i, i, vp]
    L.chz_batch_info.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_longlong)]
    L.chz_batch_destroy.argtypes = [vp]
    if hasattr(L, "chz_batch_refine_counts"):
        L.chz_batch_refine_counts.argtypes = [vp, ctypes.POINTER(i), i]


def header_symbols():
    """Function names declared in include/*.h with C linkage (chz_*)."""
    import re
    names = []
    with open(os.path.join(INCLUDE, "cuhelmholtz.h")) as f:
        for m in re.finditer(r"\b(chz_[a-z0-9_]+)\s*\(", f.read()):
            if m.group(1) not in names:
                names.append(m.group(1))
    return names


def _check(rc):
    if rc != 0:
        raise CHZError(f"status {rc}: {lib().chz_last_error().decode()}")


def _stream(t):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _vol_dims(x):
    if x.dim() != 3:
        raise ValueError("expected a padded volume shaped (NZ+2, NY+2, NX+2)")
    NZ, NY, NX = (int(s) - 2 for s in x.shape)
    return NX, NY, NZ


def do_transform(d_data, d_out=None, kinds=(DCT_2, DCT_2, DCT_2)):
    """3D transform of a padded CUDA tensor shaped (NZ+2, NY+2, NX+2), x fastest.

    d_out None (or d_data): in place.  Otherwise the result is written into
    d_out.  float64 -> FP64 path, float32 -> FP32 path.  Returns the result tensor.
    """
    import torch
    if not d_data.is_cuda or not d_data.is_contiguous():
        raise ValueError("d_data must be a contiguous CUDA tensor")
    NX, NY, NZ = _vol_dims(d_data)
    out = d_data if d_out is None else d_out
    if out.shape != d_data.shape or out.dtype != d_data.dtype or not out.is_contiguous():
        raise ValueError("d_out must match d_data")
    fn = {torch.float64: lib().chz_do_transform, torch.float32: lib().chz_do_transform_f32}.get(d_data.dtype)
    if fn is None:
        raise ValueError("float64 or float32 only")
    _check(fn(ctypes.c_void_p(d_data.data_ptr()), ctypes.c_void_p(out.data_ptr()), NX, NY, NZ,
              int(kinds[0]), int(kinds[1]), int(kinds[2]), _stream(d_data)))
    return out


def _faces(bcs, n, bc, dev):
    """Validate / convert the six Neumann face arrays (None where unused)."""
    import torch
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    out, keep = [], []
    for i, b in enumerate(bcs):
        a = i // 2
        need = bc[a] in ((3, 4) if i % 2 == 0 else (2, 3))
        if b is None:
            if need:
                raise ValueError(f"boundary code {bc[a]} along axis {a} needs face array {i}")
            out.append(None)
            continue
        t = b if isinstance(b, torch.Tensor) else torch.as_tensor(b)
        t = t.to(dev, torch.float64).contiguous().reshape(-1)
        if t.numel() != sizes[i]:
            raise ValueError(f"face array {i} must have {sizes[i]} values")
        keep.append(t)
        out.append(ctypes.c_void_p(t.data_ptr()))
    return out, keep


def helmholtz_solve(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3, stencil=7):
    """In-place FFT-diagonalised solve of laplace u + lam u = f on a CUDA float64
    node tensor f shaped (NZ+1, NY+1, NX+1) (cuda_helmholtz_solver semantics).
    bcs = (bcl, bcr, bcb, bct, bcf, bce) Neumann face arrays or None.
    stencil=19: the compact 19-point operator (chz_helmholtz_solve19)."""
    import torch
    NX, NY, NZ = (int(v) for v in n)
    if not f.is_cuda or f.dtype != torch.float64 or not f.is_contiguous():
        raise ValueError("f must be a contiguous CUDA float64 tensor")
    if tuple(f.shape) != (NZ + 1, NY + 1, NX + 1):
        raise ValueError("f must be shaped (NZ+1, NY+1, NX+1)")
    faces, keep = _faces(bcs, (NX, NY, NZ), bc, f.device)
    args = []
    for a, N in enumerate((NX, NY, NZ)):
        args += [float(dom[2 * a]), float(dom[2 * a + 1]), N, int(bc[a]), faces[2 * a], faces[2 * a + 1]]
    if stencil not in (7, 19):
        raise ValueError("stencil must be 7 or 19")
    fn = lib().chz_helmholtz_solve if stencil == 7 else lib().chz_helmholtz_solve19
    _check(fn(*args, float(lam), ctypes.c_void_p(f.data_ptr()), _stream(f)))
    del keep
    return f


def hw3crt(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3, ldimf=None, mdimf=None):
    """hw3crt_ (include/hw3crt_wrapper.h; the reference's cudahelmholtz/hw3crt_wrapper.h:1-9)
    on host numpy arrays, solved on the GPU.  f: float64 array F(LDIMF, MDIMF, N+1) in
    FISHPACK order, i.e. numpy shape (NZ+1, MDIMF, LDIMF) (C order), updated in place.
    bcs = (bdxs, bdxf, bdys, bdyf, bdzs, bdzf) face arrays in FISHPACK's layouts
    ((N+1, MDIMF), (N+1, LDIMF), (M+1, LDIMF)) or None.  Returns (pertrb, ierror)."""
    import numpy as np
    L = lib()
    NX, NY, NZ = (int(v) for v in n)
    ld = NX + 1 if ldimf is None else int(ldimf)
    md = NY + 1 if mdimf is None else int(mdimf)
    if f.dtype != np.float64 or not f.flags["C_CONTIGUOUS"] or f.size < ld * md * (NZ + 1):
        raise ValueError("f must be a C-contiguous float64 array of LDIMF*MDIMF*(NZ+1) values")
    dp = ctypes.POINTER(ctypes.c_double)
    keep = []

    def arr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(dp)
    D, I = ctypes.c_double, ctypes.c_int
    xs, xf, ys, yf, zs, zf = (D(float(v)) for v in dom)
    l, m, nn = I(NX), I(NY), I(NZ)
    lb, mb, nb = (I(int(b)) for b in bc)
    el, ldi, mdi = D(float(lam)), I(ld), I(md)
    pert, ierr = D(0.0), I(-1)
    fb = [arr(b) for b in bcs]
    R = ctypes.byref
    L.hw3crt_(R(xs), R(xf), R(l), R(lb), fb[0], fb[1], R(ys), R(yf), R(m), R(mb), fb[2], fb[3],
              R(zs), R(zf), R(nn), R(nb), fb[4], fb[5], R(el), R(ldi), R(mdi), f.ctypes.data_as(dp),
              R(pert), R(ierr), None)
    del keep
    return pert.value, ierr.value


def helmholtz_adjoint(u, gu, n, bc, lam, dom=(0.0, 1.0) * 3, faces=False, want_lambda=True):
    """Vector-Jacobian product of helmholtz_solve (chz_helmholtz_adjoint).

    u: forward solution (needed for the lambda gradient), gu = dL/du, both
    CUDA float64 node tensors.  Returns (gf, face gradients (6-list, None
    where the codes read no face data, or all None if faces=False),
    glambda (0-d CUDA tensor or None))."""
    import torch
    NX, NY, NZ = (int(v) for v in n)
    shape = (NZ + 1, NY + 1, NX + 1)
    for t in (u, gu):
        if t is not None and (not t.is_cuda or t.dtype != torch.float64 or tuple(t.shape) != shape
                              or not t.is_contiguous()):
            raise ValueError("u and gu must be contiguous CUDA float64 tensors shaped (NZ+1, NY+1, NX+1)")
    gu = gu.contiguous()
    gf = torch.empty_like(gu)
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    gfaces = [None] * 6
    if faces:
        for i in range(6):
            a = i // 2
            if bc[a] in ((3, 4) if i % 2 == 0 else (2, 3)):
                gfaces[i] = torch.empty(sizes[i], dtype=torch.float64, device=gu.device)
    glam = torch.empty((), dtype=torch.float64, device=gu.device) if want_lambda else None
    args = []
    for a, N in enumerate((NX, NY, NZ)):
        args += [float(dom[2 * a]), float(dom[2 * a + 1]), N, int(bc[a])]
    ptr = [None if t is None else ctypes.c_void_p(t.data_ptr()) for t in gfaces]
    _check(lib().chz_helmholtz_adjoint(*args, float(lam), None if u is None else ctypes.c_void_p(u.data_ptr()),
                                       ctypes.c_void_p(gu.data_ptr()), ctypes.c_void_p(gf.data_ptr()), *ptr,
                                       None if glam is None else ctypes.c_void_p(glam.data_ptr()), _stream(gu)))
    return gf, gfaces, glam


def _batch_faces(bcs, n, bc, B, dev):
    """Face arrays of a batch of one geometry: each None, (size,) shared by
    the batch, or (B, size).  Returns per face (base pointer, per-problem byte
    stride) or None, and the tensors to keep alive."""
    import torch
    NX, NY, NZ = n
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    out, keep = [], []
    for i, b in enumerate(list(bcs) + [None] * (6 - len(bcs))):
        a = i // 2
        need = bc[a] in ((3, 4) if i % 2 == 0 else (2, 3))
        if b is None:
            if need:
                raise ValueError(f"boundary code {bc[a]} along axis {a} needs face array {i}")
            out.append(None)
            continue
        t = (b if isinstance(b, torch.Tensor) else torch.as_tensor(b)).to(dev, torch.float64).contiguous()
        if t.numel() == sizes[i]:
            out.append((t.data_ptr(), 0))
        elif t.numel() == B * sizes[i]:
            out.append((t.data_ptr(), 8 * sizes[i]))
        else:
            raise ValueError(f"face array {i} must have {sizes[i]} values or {B} x {sizes[i]}")
        keep.append(t)
    return out, keep


def helmholtz_solve_batch(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3):
    """In-place FP64 solve of a batch of problems of one geometry: f is a
    contiguous CUDA float64 tensor (B, NZ+1, NY+1, NX+1); lam one value or B
    values; bcs as helmholtz_solve, each face array (size,) shared by the
    batch or (B, size).  One batched solve (chz_batch_solve_cached: the plan
    for this geometry is made once and rebound on later calls)."""
    import torch
    NX, NY, NZ = (int(v) for v in n)
    if not f.is_cuda or f.dtype != torch.float64 or not f.is_contiguous() or f.dim() != 4:
        raise ValueError("f must be a contiguous CUDA float64 tensor (B, NZ+1, NY+1, NX+1)")
    if tuple(f.shape[1:]) != (NZ + 1, NY + 1, NX + 1):
        raise ValueError("f must be shaped (B, NZ+1, NY+1, NX+1)")
    B = int(f.shape[0])
    lams = [float(v) for v in (lam if hasattr(lam, "__len__") else [lam] * B)]
    if len(lams) != B:
        raise ValueError("lam must be one value or one per problem")
    faces, keep = _batch_faces(bcs, (NX, NY, NZ), bc, B, f.device)
    nodes = (NX + 1) * (NY + 1) * (NZ + 1)
    arr = (_Problem * B)()
    for k in range(B):
        q = arr[k]
        for i in range(6):
            q.domain[i] = float(dom[i])
            q.faces[i] = None if faces[i] is None else faces[i][0] + k * faces[i][1]
        for a, v in enumerate((NX, NY, NZ)):
            q.N[a] = v
            q.bc[a] = int(bc[a])
        q.lam = lams[k]
        q.f = f.data_ptr() + 8 * nodes * k
    with torch.cuda.device(f.device):
        _check(lib().chz_batch_solve_cached(arr, B, 64, 0, _stream(f)))
    del keep
    return f


def helmholtz_adjoint_batch(u, gu, n, bc, lam, dom=(0.0, 1.0) * 3, faces=False, want_lambda=True):
    """helmholtz_adjoint for a batch of one geometry (chz_helmholtz_adjoint_batch):
    u, gu (B, NZ+1, NY+1, NX+1) contiguous CUDA float64 tensors, lam one value
    or B values.  Returns (gf (B, ...), face gradients (6-list of (B, size),
    None where the codes read no face data, or all None if faces=False),
    glambda ((B,) tensor or None))."""
    import torch
    NX, NY, NZ = (int(v) for v in n)
    shape = (NZ + 1, NY + 1, NX + 1)
    for t in (u, gu):
        if t is not None and (not t.is_cuda or t.dtype != torch.float64 or t.dim() != 4
                              or tuple(t.shape[1:]) != shape or not t.is_contiguous()):
            raise ValueError("u and gu must be contiguous CUDA float64 tensors shaped (B, NZ+1, NY+1, NX+1)")
    B = int(gu.shape[0])
    if u is not None and int(u.shape[0]) != B:
        raise ValueError("u and gu must hold the same number of problems")
    lams = [float(v) for v in (lam if hasattr(lam, "__len__") else [lam] * B)]
    if len(lams) != B:
        raise ValueError("lam must be one value or one per problem")
    lam_h = (ctypes.c_double * B)(*lams)
    gf = torch.empty_like(gu)
    sizes = [(NZ + 1) * (NY + 1)] * 2 + [(NZ + 1) * (NX + 1)] * 2 + [(NY + 1) * (NX + 1)] * 2
    gfaces = [None] * 6
    if faces:
        for i in range(6):
            a = i // 2
            if bc[a] in ((3, 4) if i % 2 == 0 else (2, 3)):
                gfaces[i] = torch.empty((B, sizes[i]), dtype=torch.float64, device=gu.device)
    glam = torch.empty(B, dtype=torch.float64, device=gu.device) if want_lambda else None
    args = []
    for a, N in enumerate((NX, NY, NZ)):
        args += [float(dom[2 * a]), float(dom[2 * a + 1]), N, int(bc[a])]
    ptr = [None if t is None else ctypes.c_void_p(t.data_ptr()) for t in gfaces]
    with torch.cuda.device(gu.device):
        _check(lib().chz_helmholtz_adjoint_batch(*args, B, ctypes.cast(lam_h, ctypes.c_void_p),
                                                 None if u is None else ctypes.c_void_p(u.data_ptr()),
                                                 ctypes.c_void_p(gu.data_ptr()), ctypes.c_void_p(gf.data_ptr()), *ptr,
                                                 None if glam is None else ctypes.c_void_p(glam.data_ptr()),
                                                 _stream(gu)))
    return gf, gfaces, glam


def release():
    """Free the cached twiddle tables, workspaces and batched plans of the current
    device (chz_release; the reference's freeMemory_cubic)."""
    lib().chz_release()


def set_timing(on):
    """Record per-pass events on every transform / solve (reference TIME_TEST counterpart)."""
    lib().chz_set_timing(1 if on else 0)


def get_timing(max_passes=4096):
    """[(pass name, ms)] of every pass recorded since timing was switched on or
    last read (waits for the events, then clears them)."""
    ms = (ctypes.c_double * max_passes)()
    names = ctypes.create_string_buffer(16 * max_passes)
    n = lib().chz_get_timing(ms, max_passes, names, 16 * max_passes)
    if n < 0:
        _check(n)
    return list(zip(names.value.decode().split(",")[:n], [ms[i] for i in range(n)]))


class BatchSolver:
    """Batched solve of many independent problems (chz_batch_create): one
    hipGraph per plan, one launch per pass and size class.

    problems: sequence of (f, n, bc, lam[, bcs[, dom]]) with f a contiguous
    CUDA float64 node tensor shaped (NZ+1, NY+1, NX+1) -- bound to the plan:
    solve() solves whatever the tensors hold, in place.  prec 64 = FP64,
    prec 32 = FP32 transforms + `refine` FP64 refinement steps (refine < 0:
    per-problem counts up to -refine from each operator's spectral
    condition number, see refine_counts()).
    """

    def __init__(self, problems, prec=64, refine=0):
        import torch
        arr, self._keep, dev = self._problems(problems)
        self.device = dev if dev is not None else torch.device("cuda", torch.cuda.current_device())
        self.count = len(problems)
        self.prec, self.refine = prec, refine
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(lib().chz_batch_create(arr, len(problems), int(prec), int(refine), ctypes.byref(self._plan)))

    @staticmethod
    def _problems(problems):
        import torch
        keep = []
        arr = (_Problem * len(problems))()
        dev = None
        for k, pr in enumerate(problems):
            f, n, bc, lam = pr[:4]
            bcs = pr[4] if len(pr) > 4 and pr[4] is not None else (None,) * 6
            dom = pr[5] if len(pr) > 5 else (0.0, 1.0) * 3
            NX, NY, NZ = (int(v) for v in n)
            if not f.is_cuda or f.dtype != torch.float64 or not f.is_contiguous():
                raise ValueError(f"problem {k}: f must be a contiguous CUDA float64 tensor")
            if tuple(f.shape) != (NZ + 1, NY + 1, NX + 1):
                raise ValueError(f"problem {k}: f must be shaped (NZ+1, NY+1, NX+1)")
            if dev is None:
                dev = f.device
            elif f.device != dev:
                raise ValueError("all problems must live on one device")
            faces, kf = _faces(bcs, (NX, NY, NZ), bc, f.device)
            keep += [f] + kf
            q = arr[k]
            for i in range(6):
                q.domain[i] = float(dom[i])
                q.faces[i] = faces[i].value if faces[i] is not None else None
            for a, v in enumerate((NX, NY, NZ)):
                q.N[a] = v
                q.bc[a] = int(bc[a])
            q.lam = float(lam)
            q.f = f.data_ptr()
        return arr, keep, dev

    def solve(self, graph=True):
        """Run the batched solve on the current stream (graph=False: direct
        launches, timed per launch when set_timing is on)."""
        import torch
        with torch.cuda.device(self.device):
            s = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            _check(lib().chz_batch_solve(self._plan, 1 if graph else 0, s))

    def rebind(self, problems):
        """Point the plan at new problems of the same geometry (chz_batch_rebind):
        new f / face tensors and lambdas, uploaded on the current stream after
        the plan's earlier solves (no host wait)."""
        import torch
        arr, keep, _ = self._problems(problems)
        with torch.cuda.device(self.device):
            s = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
            _check(lib().chz_batch_rebind(self._plan, arr, len(problems), s))
        self._keep = keep

    def refine_counts(self):
        """Refinement steps of each problem (chz_batch_refine_counts)."""
        arr = (ctypes.c_int * self.count)()
        _check(lib().chz_batch_refine_counts(self._plan, arr, self.count))
        return list(arr)

    def info(self):
        n, w, t = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().chz_batch_info(self._plan, ctypes.byref(n), ctypes.byref(w), ctypes.byref(t)))
        return {"launches": n.value, "work_bytes": w.value, "table_bytes": t.value}

    def close(self):
        if self._plan:
            import torch
            torch.cuda.synchronize(self.device)
            lib().chz_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p()
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# roles of the launch configurations (ROLE_ZMIDN: the fused z pass of a
# solve with non-periodic z; a size without its own uses the ROLE_ZMID one)
ROLE_ROWS, ROLE_COLS, ROLE_XFWD, ROLE_ZMID, ROLE_ZCOLS, ROLE_ZMIDN = range(6)
ROLE_NAMES = ["rows", "cols", "xfwd", "zmid", "zcols", "zmidn"]


def helmholtz_solve_mixed(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3, refine=1):
    """In-place solve with FP32 transforms / divide and `refine` FP64 iterative
    refinement steps (same arguments as helmholtz_solve)."""
    import torch
    NX, NY, NZ = (int(v) for v in n)
    if not f.is_cuda or f.dtype != torch.float64 or not f.is_contiguous():
        raise ValueError("f must be a contiguous CUDA float64 tensor")
    if tuple(f.shape) != (NZ + 1, NY + 1, NX + 1):
        raise ValueError("f must be shaped (NZ+1, NY+1, NX+1)")
    faces, keep = _faces(bcs, (NX, NY, NZ), bc, f.device)
    args = []
    for a, N in enumerate((NX, NY, NZ)):
        args += [float(dom[2 * a]), float(dom[2 * a + 1]), N, int(bc[a]), faces[2 * a], faces[2 * a + 1]]
    _check(lib().chz_helmholtz_solve_mixed(*args, float(lam), ctypes.c_void_p(f.data_ptr()), int(refine), _stream(f)))
    del keep
    return f


def helmholtz_solve_mixed_fp32only(f, n, bc, lam, bcs=(None,) * 6, dom=(0.0, 1.0) * 3):
    """FP32 solve without refinement (the building block the autotuner times)."""
    return helmholtz_solve_mixed(f, n, bc, lam, bcs, dom, refine=0)


def axis_config(N, prec=64, rows=True, role=None):
    """Launch configuration of a pass kernel (role defaults to rows/cols)."""
    if role is None:
        role = ROLE_ROWS if rows else ROLE_COLS
    e, t, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = ctypes.c_ulonglong()
    _check(lib().chz_axis_config(N, prec, role, ctypes.byref(e), ctypes.byref(t), ctypes.byref(l),
                                 ctypes.byref(p)))
    radices = []
    v = p.value
    while v:
        radices.append(v & 255)
        v >>= 8
    return {"ept": e.value, "tpl": t.value, "lines": l.value, "radices": radices}


def set_axis_config(N, prec, role, ept, tpl, lines):
    """Pin a launch configuration; raises CHZError if it cannot run."""
    _check(lib().chz_set_axis_config(N, prec, role, ept, tpl, lines))


def reset_axis_configs():
    """Forget every pinned or loaded configuration (back to the built-in choice)."""
    lib().chz_reset_axis_configs()


def tuning_table_path(arch="gfx950"):
    """The tuning table the library loads at start-up (beside the library)."""
    return os.path.join(HERE, "tuning", f"{arch}.txt")


def load_tuning(path=None):
    """Pin every configuration of a tuning table (default: the shipped one);
    returns how many rows were valid."""
    n = lib().chz_load_tuning((path or tuning_table_path()).encode())
    _check(min(n, 0))
    return n


def radix_plan(M, ept, tpl):
    """The library's radix plan, one byte per stage radix (first stage in the
    low byte; a 1 is a direct stage of the remaining factor), for an M-point
    FFT with at most ept points per thread and tpl threads per line, or None
    if the shape cannot run (host only)."""
    p = ctypes.c_ulonglong()
    if lib().chz_radix_plan(int(M), int(ept), int(tpl), ctypes.byref(p)) != 0:
        return None
    return p.value


def compiled_kernels():
    """The compiled fixed-shape pass kernels as tuples (form, prec, op, N, ept,
    tpl, lines, plan, ks), form 0 single-problem lean tile I/O, 1 its batched
    form, 2 single-problem fixed shape, 3 batched fixed shape; ks the kind
    selector (-1 every kind, else the pass kind or the fused z pass's boundary
    code) (host only)."""
    L = lib()
    n = L.chz_compiled_kernels(None, 0)
    buf = (ctypes.c_int * (10 * n))()
    L.chz_compiled_kernels(buf, n)
    out = []
    for k in range(n):
        r = buf[10 * k:10 * k + 10]
        out.append(tuple(r[:7]) + (((r[8] & 0xffffffff) << 32) | (r[7] & 0xffffffff), r[9]))
    return out


def in_window(kind, N):
    return {DST_1: (1, N - 1), DCT_1: (0, N), DCT_2: (0, N - 1), DST_2: (1, N), DCT_3: (0, N - 1),
            DST_3: (1, N), DFT_R2C: (0, N - 1), DFT_C2R: (0, N + 1)}[kind]


def out_window(kind, N):
    return {DST_1: (1, N - 1), DCT_1: (0, N), DCT_2: (0, N - 1), DST_2: (1, N), DCT_3: (0, N - 1),
            DST_3: (1, N), DFT_R2C: (0, N + 1), DFT_C2R: (0, N - 1)}[kind]
