"""ctypes binding of libgimic_hip.so (include/gimic_hip.h).

The library is the only compute path: importing this module fails loudly
if the HIP library has not been built (python -m gimic_amd.build).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libgimic_hip.so")

SPIN = {"total": 0, "alpha": 1, "beta": 2, "spindens": 3}
PREC = {"fp64": 0, "mixed": 1, "bf16x6": 1}

EXPORTS = [
    "gimic_init", "gimic_finalize", "gimic_set_uhf", "gimic_set_magnet", "gimic_set_spin",
    "gimic_set_screening", "gimic_calc_jtensor", "gimic_calc_jvector", "gimic_calc_modj",
    "gimic_get_gauss_points", "mkgausspoints", "gimic_last_error", "gimic_ctx_create_files",
    "gimic_ctx_create", "gimic_ctx_set_density", "gimic_ctx_set_flags", "gimic_ctx_info",
    "gimic_ctx_destroy", "gimic_ctx_atoms", "gimic_ctx_atom_symbols", "gimic_ctx_set_timing", "gimic_ctx_get_timing", "gimic_ctx_tuning", "gimic_ctx_tuning_log", "gimic_ctx_set_mixed_kernel", "gimic_ctx_set_tile_stats", "gimic_ctx_tile_stats", "gimic_ctx_set_fp64_radius", "gimic_ctx_fp64_refined", "gimic_ctx_set_chunk_bytes", "gimic_ctx_chunk_bytes", "gimic_ctx_set_graph", "gimic_ctx_graph_stats", "gimic_jtensor_batch", "gimic_jtensor_batch_device", "gimic_jtensor_batch_device_noalloc",
    "gimic_jtensor_multi_device", "gimic_flux_planes", "gimic_flux_plane_grids", "gimic_ctx_properties", "gimic_sph_transform", "gimic_quadrature",
    "gimic_selftest_mfma_layout", "gimic_selftest_mfma32_layout", "gimic_mlp_create", "gimic_mlp_destroy", "gimic_mlp_set_stream", "gimic_mlp_info",
    "gimic_mlp_params", "gimic_mlp_loss_grad", "gimic_mlp_train_step", "gimic_mlp_predict",
    "gimic_mlp_predict_unfused",
]

_lib = None


class GimicError(RuntimeError):
    pass


def load():
    """Load the library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise GimicError(f"{LIBPATH} not found: build it with `python -m gimic_amd.build`")
    # PyTorch's ROCm wheels bundle their own libamdhip64.so.7 and
    # libhsa-runtime64.so.  Loading this library first would bind the system
    # HIP/HSA runtime, and a later `import torch` would bring a second HSA
    # runtime into the process, which then finds no device.  Importing torch
    # first makes this library bind torch's runtime (same soname), so the
    # process holds one HIP runtime and device pointers pass between the two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIBPATH)
    P, I, D, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
    L.gimic_last_error.restype = ctypes.c_char_p
    sig = {
        "gimic_init": [ctypes.c_char_p, ctypes.c_char_p],
        "gimic_finalize": [],
        "gimic_set_uhf": [P],
        "gimic_set_magnet": [P],
        "gimic_set_spin": [ctypes.c_char_p],
        "gimic_set_screening": [P],
        "gimic_calc_jtensor": [P, P],
        "gimic_calc_jvector": [P, P],
        "gimic_calc_modj": [P, P],
        "gimic_get_gauss_points": [P, P, P, P, P, P],
        "mkgausspoints": [P, P, P, P, P, P],
        "gimic_ctx_create_files": [P, ctypes.c_char_p, ctypes.c_char_p, I, D],
        "gimic_ctx_create": [P, I, P, I, P, P, P, P, P, P, I, I, D],
        "gimic_ctx_set_density": [P, I, P],
        "gimic_ctx_set_flags": [P, I, I, I],
        "gimic_ctx_info": [P, P, P, P, P, P],
        "gimic_ctx_destroy": [P],
        "gimic_ctx_atoms": [P, P, P],
        "gimic_ctx_atom_symbols": [P, P],
        "gimic_ctx_set_timing": [P, I],
        "gimic_ctx_get_timing": [P, P, P],
        "gimic_ctx_tuning": [P, P, P, P],
        "gimic_ctx_tuning_log": [P, P, P, P],
        "gimic_ctx_set_mixed_kernel": [P, I, I],
        "gimic_ctx_set_tile_stats": [P, I],
        "gimic_ctx_set_fp64_radius": [P, D],
        "gimic_ctx_fp64_refined": [P, P, P],
        "gimic_ctx_set_chunk_bytes": [P, I64],
        "gimic_ctx_chunk_bytes": [P, P],
        "gimic_ctx_tile_stats": [P, P],
        "gimic_ctx_set_graph": [P, I],
        "gimic_ctx_graph_stats": [P, P, P],
        "gimic_jtensor_batch": [P, I64, P, P, I, I],
        "gimic_jtensor_batch_device": [P, I64, P, P, I, I, P],
        "gimic_jtensor_batch_device_noalloc": [P, I64, P, P, I, I, P],
        "gimic_jtensor_multi_device": [P, I64, P, P, I, P, I, P],
        "gimic_flux_planes": [P, I, P, P, P, P, P, P, P, I, I, P, P, P, P],
        "gimic_flux_plane_grids": [P, I, P, P, P, P, P, P, P, P, P, P, P, P, P, I, I, P, P, P],
        "gimic_ctx_properties": [P, I64, P, P, I, P, I, P, I, I, P, P, P],
        "gimic_sph_transform": [I, P],
        "gimic_quadrature": [D, D, I, I, I, P, P],
        "gimic_selftest_mfma_layout": [P, P],
        "gimic_selftest_mfma32_layout": [P],
        "gimic_mlp_create": [P, D, P, P, ctypes.c_uint64],
        "gimic_mlp_destroy": [P],
        "gimic_mlp_set_stream": [P, P],
        "gimic_mlp_info": [P, P, P],
        "gimic_mlp_params": [P, P, I],
        "gimic_mlp_loss_grad": [P, ctypes.c_longlong, P, P, D, P, P],
        "gimic_mlp_train_step": [P, ctypes.c_longlong, P, P, D, D, P],
        "gimic_mlp_predict": [P, ctypes.c_longlong, P, P, P],
        "gimic_mlp_predict_unfused": [P, ctypes.c_longlong, P, P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        if name.startswith(("gimic_ctx", "gimic_mlp")) or name in ("gimic_jtensor_batch", "gimic_jtensor_batch_device",
                                                    "gimic_jtensor_batch_device_noalloc", "gimic_jtensor_batch_device_noalloc",
                                                    "gimic_jtensor_multi_device", "gimic_flux_planes",
                                                    "gimic_flux_plane_grids",
                                                    "gimic_sph_transform", "gimic_quadrature",
                                                    "gimic_selftest_mfma_layout",
                                                    "gimic_selftest_mfma32_layout"):
            fn.restype = I
        else:
            fn.restype = None
    _lib = L
    return L


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def check(rc):
    if rc != 0:
        msg = load().gimic_last_error().decode(errors="replace")
        raise GimicError(f"gimic error {rc}: {msg}")


def _f64(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a if shape is None else a.reshape(shape)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def sph_transform(l):
    nc = (l + 1) * (l + 2) // 2
    T = np.zeros((2 * l + 1, nc))
    check(load().gimic_sph_transform(l, ptr(T)))
    return T


def quadrature(a, b, order, npts, kind="gauss"):
    """Composite nodes/weights as the reference's setup_gauss_data."""
    pts = np.zeros(npts)
    w = np.zeros(npts)
    check(load().gimic_quadrature(float(a), float(b), int(order), int(npts),
                                  {"gauss": 0, "lobatto": 1}[kind], ptr(pts), ptr(w)))
    return pts, w


def plane_grid_descriptors(grids, bfields):
    """Per-plane arrays of gimic_flux_plane_grids for product grids
    (gimic_amd.grid.Grid): n1, n2 (int32), origins (n,3), axes (n,9: v1 v2
    v3), p3w3 (n,2), the concatenated 1-D rules nodes1/wts1/nodes2/wts2,
    normals (v3), bfields, centers (Grid.center(), the same operations in
    the same order), radii, and owner (the grid of each plane).  A grid with
    n3 > 1 layers gives n3 planes.  Vectorised over the grids: a scan of a
    few hundred planes costs well under a millisecond here."""
    if len(grids) == 0 or len(grids) != len(bfields):
        raise ValueError("need one field per grid and at least one grid")
    if any(g.mode == "file" for g in grids):
        raise ValueError("flux_plane_grids needs product grids, not file grids")
    G = len(grids)
    p1s, p2s, p3s = [g.pts[0] for g in grids], [g.pts[1] for g in grids], [g.pts[2] for g in grids]
    w1s, w2s, w3s = [g.wgt[0] for g in grids], [g.wgt[1] for g in grids], [g.wgt[2] for g in grids]
    n3 = np.fromiter(map(len, p3s), dtype=np.int64, count=G)
    layered = bool((n3 != 1).any())
    owner = np.repeat(np.arange(G), n3) if layered else np.arange(G)
    basv = np.array([g.basv for g in grids], dtype=np.float64)  # columns v1 v2 v3
    org = np.array([g.origin for g in grids], dtype=np.float64)
    v1, v2, v3 = basv[:, :, 0], basv[:, :, 1], basv[:, :, 2]
    col = lambda vals: np.array(vals, dtype=np.float64)[:, None]
    # Grid.center(): (gridpoint(n1, 1, 1) + gridpoint(1, n2, 1)) * 0.5
    p3f = col([p[0] for p in p3s])
    ca = ((org + col([p[-1] for p in p1s]) * v1) + col([p[0] for p in p2s]) * v2) + p3f * v3
    cb = ((org + col([p[0] for p in p1s]) * v1) + col([p[-1] for p in p2s]) * v2) + p3f * v3
    ctr = (ca + cb) * 0.5
    axes = basv.transpose(0, 2, 1).reshape(-1, 9)
    n1 = np.fromiter(map(len, p1s), dtype=np.int32, count=G)
    n2 = np.fromiter(map(len, p2s), dtype=np.int32, count=G)
    rad = np.fromiter((float(g.radius) for g in grids), dtype=np.float64, count=G)
    bf = np.asarray(bfields, dtype=np.float64).reshape(-1, 3)
    cat = lambda parts: np.ascontiguousarray(np.concatenate(parts), dtype=np.float64)
    per = (lambda parts: [parts[i] for i in owner.tolist()]) if layered else (lambda parts: parts)
    pick = (lambda a: np.ascontiguousarray(a[owner])) if layered else np.ascontiguousarray
    return dict(
        n1=pick(n1), n2=pick(n2), origins=pick(org), axes=pick(axes),
        p3w3=np.ascontiguousarray(np.stack([cat(p3s), cat(w3s)], 1)),
        nodes1=cat(per(p1s)), wts1=cat(per(w1s)), nodes2=cat(per(p2s)), wts2=cat(per(w2s)),
        normals=pick(v3), bfields=pick(bf), centers=pick(ctr), radii=pick(rad), owner=owner)


class Context:
    """Handle of one basis + density family resident on the current GPU."""

    def __init__(self, handle):
        self._h = handle
        self._lib = load()

    # -- construction ----------------------------------------------------
    @classmethod
    def from_files(cls, molfile, densfile=None, *, uhf=False, screening=1e-8):
        L = load()
        h = ctypes.c_void_p()
        check(L.gimic_ctx_create_files(ctypes.byref(h), os.fsencode(molfile),
                                       os.fsencode(densfile) if densfile else None,
                                       int(bool(uhf)), float(screening if screening else -1.0)))
        return cls(h)

    @classmethod
    def from_arrays(cls, coords, shell_atom, shell_l, shell_nprim, shell_poff, exps, coefs, *,
                    turbomole=False, spherical=False, screening=1e-8):
        L = load()
        h = ctypes.c_void_p()
        coords = _f64(coords, (-1, 3))
        sa, sl, sn, so = _i32(shell_atom), _i32(shell_l), _i32(shell_nprim), _i32(shell_poff)
        ex, co = _f64(exps), _f64(coefs)
        check(L.gimic_ctx_create(ctypes.byref(h), coords.shape[0], ptr(coords), len(sl), ptr(sa),
                                 ptr(sl), ptr(sn), ptr(so), ptr(ex), ptr(co), int(bool(turbomole)),
                                 int(bool(spherical)), float(screening if screening else -1.0)))
        return cls(h)

    def close(self):
        if self._h is not None:
            check(self._lib.gimic_ctx_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------
    def info(self):
        v = [ctypes.c_int() for _ in range(5)]
        check(self._lib.gimic_ctx_info(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["nbf", "nbf_cart", "natoms", "nshell", "nset"], [x.value for x in v]))

    @property
    def nbf(self):
        return self.info()["nbf"]

    def set_timing(self, on=True):
        check(self._lib.gimic_ctx_set_timing(self._h, int(bool(on))))

    def get_timing(self):
        """{kernel: (total ms, launches)} for basis_eval, contract, finalize"""
        ms = np.zeros(3)
        n = np.zeros(3, dtype=np.int64)
        check(self._lib.gimic_ctx_get_timing(self._h, ptr(ms), ptr(n)))
        return {k: (float(ms[i]), int(n[i]))
                for i, k in enumerate(["basis_eval", "contract", "finalize"])}

    def tuning(self):
        """(chosen mixed kernel variant, its N-split count, [best ms of
        variants 1 to 9, 0 where not timed])"""
        v, sp = ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(9)
        check(self._lib.gimic_ctx_tuning(self._h, ctypes.byref(v), ctypes.byref(sp), ptr(ms)))
        return v.value, sp.value, ms.tolist()

    def tuning_log(self):
        """[(variant, split, ms)] of every candidate the tuner timed"""
        n = ctypes.c_int(32)
        codes = np.zeros(32, dtype=np.int32)
        ms = np.zeros(32)
        check(self._lib.gimic_ctx_tuning_log(self._h, ctypes.byref(n), ptr(codes), ptr(ms)))
        return [(int(c) // 100, int(c) % 100, float(t)) for c, t in zip(codes[:n.value], ms)]

    def set_mixed_kernel(self, variant=0, split=0):
        """fix the mixed contraction kernel (3/4: 4 wavefronts, 128-row tiles,
        double/triple buffered; 5: 4 wavefronts, 256-row tiles; 8/9: 8
        wavefronts, 256-row tiles, double/triple buffered; `split` N-ranges
        per row tile, 0 = by workgroup count) or 0 for automatic selection"""
        check(self._lib.gimic_ctx_set_mixed_kernel(self._h, int(variant), int(split)))

    def set_fp64_radius(self, radius):
        """mixed precision: points within `radius` bohr of a nucleus are
        evaluated again in FP64 (default 0.4; 0 turns it off)"""
        check(self._lib.gimic_ctx_set_fp64_radius(self._h, float(radius)))

    def set_chunk_bytes(self, nbytes):
        """device memory for the basis buffers of one chunk of points (0: the
        default, 8 GiB; -1: a third of the device's memory, at most 96 GiB)"""
        check(self._lib.gimic_ctx_set_chunk_bytes(self._h, int(nbytes)))

    def chunk_bytes(self):
        """the chunk memory in force, in bytes"""
        n = ctypes.c_int64(0)
        check(self._lib.gimic_ctx_chunk_bytes(self._h, ctypes.byref(n)))
        return n.value

    def fp64_refined(self):
        """(radius, points the last mixed call evaluated in FP64)"""
        r = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        check(self._lib.gimic_ctx_fp64_refined(self._h, ctypes.byref(r), ctypes.byref(n)))
        return r.value, int(n.value)

    def set_tile_stats(self, on=True):
        """collect screening block-list counts of eager evaluations (each
        copies its counts to the host: keep it off in timed runs)"""
        check(self._lib.gimic_ctx_set_tile_stats(self._h, 1 if on else 0))

    def tile_stats(self):
        """dict(tiles, blocks, blocks_sq, tile_rows, nblk, blocks_kn,
        groups_sq): row tiles seen, sum of listed blocks and of their
        squares, rows per tile, blocks of a dense row, the sum of listed
        blocks x live 8-function groups / 4 (the FP64 kernel's block
        products) and of squared live groups (the pre-split mixed kernels'
        group products)"""
        out = np.zeros(7)
        check(self._lib.gimic_ctx_tile_stats(self._h, ptr(out)))
        return dict(tiles=int(out[0]), blocks=float(out[1]), blocks_sq=float(out[2]),
                    tile_rows=int(out[3]), nblk=int(out[4]), blocks_kn=float(out[5]),
                    groups_sq=float(out[6]))

    def set_graph(self, enable=True):
        """capture flux_planes batches into one reusable hipGraph launch"""
        check(self._lib.gimic_ctx_set_graph(self._h, 1 if enable else 0))

    def graph_stats(self):
        """(graph builds, graph launches) of flux_planes in graph mode"""
        b, n = ctypes.c_longlong(), ctypes.c_longlong()
        check(self._lib.gimic_ctx_graph_stats(self._h, ctypes.byref(b), ctypes.byref(n)))
        return b.value, n.value

    def atoms(self):
        """(coords (natoms,3) bohr, charges (natoms,))"""
        n = self.info()["natoms"]
        c = np.zeros((n, 3))
        q = np.zeros(n)
        check(self._lib.gimic_ctx_atoms(self._h, ptr(c), ptr(q)))
        return c, q

    def symbols(self):
        """two-character atom symbols as the MOL file gives them"""
        n = self.info()["natoms"]
        buf = ctypes.create_string_buffer(2 * n)
        check(self._lib.gimic_ctx_atom_symbols(self._h, buf))
        raw = buf.raw.decode(errors="replace")
        return [raw[2 * a:2 * a + 2] for a in range(n)]

    def set_density(self, dens):
        """dens: (nset, 4, n, n) with [.., mu, nu] = D(mu, nu)."""
        d = np.asarray(dens, dtype=np.float64)
        if d.ndim == 3:
            d = d[None]
        nset = d.shape[0]
        flat = np.ascontiguousarray(d.transpose(0, 1, 3, 2)).reshape(-1)
        check(self._lib.gimic_ctx_set_density(self._h, nset, ptr(flat)))

    def set_flags(self, giao=True, diamag=True, paramag=True):
        check(self._lib.gimic_ctx_set_flags(self._h, int(giao), int(diamag), int(paramag)))

    # -- evaluation --------------------------------------------------------
    def jtensor(self, points, spin="total", precision="fp64"):
        """(N,3) host points -> (N,9) tensors, element m + 3b = J(m,b)."""
        pts = _f64(points, (-1, 3))
        out = np.zeros((pts.shape[0], 9))
        check(self._lib.gimic_jtensor_batch(self._h, pts.shape[0], ptr(pts), ptr(out),
                                            SPIN[spin], PREC[precision]))
        return out

    def jtensor_device(self, points, spin="total", precision="fp64", out=None, stream=None):
        """torch float64 (N,3) tensor on the GPU -> (N,9) tensor (asynchronous)."""
        import torch
        pts = points.contiguous()
        assert pts.dtype == torch.float64 and pts.is_cuda
        n = pts.shape[0]
        if out is None:
            out = torch.empty((n, 9), dtype=torch.float64, device=pts.device)
        s = ctypes.c_void_p(stream) if stream is not None else \
            ctypes.c_void_p(torch.cuda.current_stream(pts.device).cuda_stream)
        check(self._lib.gimic_jtensor_batch_device(self._h, n, ctypes.c_void_p(pts.data_ptr()),
                                                   ctypes.c_void_p(out.data_ptr()), SPIN[spin],
                                                   PREC[precision], s))
        return out

    def jtensor_device_noalloc(self, points, spin="total", precision="fp64", out=None,
                               stream=None):
        """as jtensor_device through the allocation-free entry point (work
        buffers in the library's per-device scratch array; the right operands
        must be resident from an earlier jtensor_device call)"""
        import torch
        pts = points.contiguous()
        assert pts.dtype == torch.float64 and pts.is_cuda
        n = pts.shape[0]
        if out is None:
            out = torch.empty((n, 9), dtype=torch.float64, device=pts.device)
        s = ctypes.c_void_p(stream) if stream is not None else \
            ctypes.c_void_p(torch.cuda.current_stream(pts.device).cuda_stream)
        check(self._lib.gimic_jtensor_batch_device_noalloc(
            self._h, n, ctypes.c_void_p(pts.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            SPIN[spin], PREC[precision], s))
        return out

    def jtensor_multi_device(self, points, spins, precision="fp64", out=None, stream=None):
        import torch
        pts = points.contiguous()
        n = pts.shape[0]
        sp = _i32([SPIN[s] for s in spins])
        if out is None:
            out = torch.empty((len(spins), n, 9), dtype=torch.float64, device=pts.device)
        s = ctypes.c_void_p(stream) if stream is not None else \
            ctypes.c_void_p(torch.cuda.current_stream(pts.device).cuda_stream)
        check(self._lib.gimic_jtensor_multi_device(self._h, n, ctypes.c_void_p(pts.data_ptr()),
                                                   ctypes.c_void_p(out.data_ptr()), len(spins),
                                                   ptr(sp), PREC[precision], s))
        return out

    def properties(self, points, weights, part_sizes, nuclei, spin="total", precision="fp64",
                   return_jt=False):
        """Shielding and magnetizability quadrature on an external grid
        (jfield.f90:584-929): points (N,3) with weights (N), split into
        consecutive partitions of part_sizes points (one per atom grid),
        nuclei (K,3).  Returns dict(shield=(K, P, 5), magnet=(P, 5)), q =
        weighted sums of the xx, yy, zz integrands, positive and negative
        parts of their sum (shielding in ppm, magnetizability in au), and
        the tensors if return_jt."""
        r = _f64(points).reshape(-1, 3)
        w = _f64(weights).reshape(-1)
        sizes = np.asarray(part_sizes, dtype=np.int64).reshape(-1)
        offs = np.zeros(len(sizes) + 1, dtype=np.int64)
        offs[1:] = np.cumsum(sizes)
        nuc = _f64(nuclei).reshape(-1, 3)
        shield = np.zeros((len(nuc), len(sizes), 5))
        magnet = np.zeros((len(sizes), 5))
        jt = np.zeros((len(r), 9)) if return_jt else None
        check(self._lib.gimic_ctx_properties(self._h, len(r), ptr(r), ptr(w), len(sizes), ptr(offs),
                                             len(nuc), ptr(nuc), SPIN[spin], PREC[precision],
                                             ptr(shield), ptr(magnet),
                                             ptr(jt) if jt is not None else None))
        out = dict(shield=shield, magnet=magnet)
        if return_jt:
            out["jt"] = jt
        return out

    def flux_planes(self, planes, spin="total", precision="fp64", return_jt=False,
                    modulus=False, acid=False):
        """planes: list of dicts with points (N,3), weights (N), normal, bfield,
        center, radius.  Returns (nplanes, 3): net, positive, negative (au);
        with modulus=True also the signed-modulus sums (nplanes, 3), with
        acid=True the ACID integrals sum w dT^2 (nplanes,), then the tensors
        if return_jt."""
        offs = np.zeros(len(planes) + 1, dtype=np.int64)
        for k, p in enumerate(planes):
            offs[k + 1] = offs[k] + len(p["weights"])
        r = _f64(np.concatenate([np.asarray(p["points"]).reshape(-1, 3) for p in planes]))
        w = _f64(np.concatenate([np.asarray(p["weights"]).reshape(-1) for p in planes]))
        nrm = _f64([p["normal"] for p in planes])
        bf = _f64([p["bfield"] for p in planes])
        ctr = _f64([p["center"] for p in planes])
        rad = _f64([p.get("radius", 1e10) for p in planes])
        out = np.zeros((len(planes), 3))
        mod = np.zeros((len(planes), 3)) if modulus else None
        aci = np.zeros(len(planes)) if acid else None
        jt = np.zeros((int(offs[-1]), 9)) if return_jt else None
        check(self._lib.gimic_flux_planes(self._h, len(planes), ptr(offs), ptr(r), ptr(w),
                                          ptr(nrm), ptr(bf), ptr(ctr), ptr(rad), SPIN[spin],
                                          PREC[precision], ptr(out),
                                          ptr(mod) if mod is not None else None,
                                          ptr(aci) if aci is not None else None,
                                          ptr(jt) if jt is not None else None))
        res = (out,) + ((mod,) if modulus else ()) + ((aci,) if acid else ()) + \
            ((jt,) if return_jt else ())
        return res if len(res) > 1 else out

    def flux_plane_grids(self, grids, bfields, spin="total", precision="fp64", modulus=False,
                         acid=False):
        """Flux through product-quadrature grids (gimic_amd.grid.Grid, not file
        grids) whose points and weights the device generates from the 1-D rules:
        only O(n1 + n2) numbers per grid cross to the device.  Same results, bit
        for bit, as flux_planes on [flux.plane_of(g, b) ...].  A grid with n3 > 1
        layers goes in as n3 planes whose sums are added (all three sums are
        additive over points)."""
        d = plane_grid_descriptors(grids, bfields)
        npl = len(d["n1"])
        out = np.zeros((npl, 3))
        mod = np.zeros((npl, 3)) if modulus else None
        aci = np.zeros(npl) if acid else None
        n1a, n2a = d["n1"], d["n2"]
        arrs = [d[k] for k in ("origins", "axes", "p3w3", "nodes1", "wts1", "nodes2", "wts2",
                               "normals", "bfields", "centers", "radii")]
        owner = d["owner"]
        check(self._lib.gimic_flux_plane_grids(self._h, npl, ptr(n1a), ptr(n2a),
                                               *[ptr(a) for a in arrs], SPIN[spin],
                                               PREC[precision], ptr(out),
                                               ptr(mod) if mod is not None else None,
                                               ptr(aci) if aci is not None else None))
        ng = len(grids)
        fold = lambda a: np.stack([a[owner == i].sum(axis=0) for i in range(ng)]) \
            if npl != ng else a
        res = (fold(out),) + ((fold(mod),) if modulus else ()) + ((fold(aci),) if acid else ())
        return res if len(res) > 1 else res[0]
