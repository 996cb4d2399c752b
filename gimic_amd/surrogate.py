"""Learned surrogate of the J-tensor field for dense visualisation grids.

A per-molecule coordinate network (the library's gimic_mlp_* entry points:
Fourier features of r, 4 SiLU layers of width 128, 9 outputs; trained and run
by the library's own HIP kernels, BF16 operands with FP32 master weights and
Adam) is fitted to tensors that the exact kernel computes.  The exact kernel
stays in charge:

* every point within `exclusion` bohr of a nucleus is computed exactly;
* the bounding box is divided into ncell^3 cells; 1% of the samples are held
  out and a cell is served by the network only when its held-out relative
  error sqrt(sum |J_net - J|^2 / sum |J|^2) is at most `tol` and it holds
  at least `min_holdout` held-out points; every other cell is exact;
* `audit()` evaluates a sample of the served points with the exact kernel and
  reports the error actually delivered.

The surrogate has no counterpart in the reference; its input is the same
molecule/density context as calc_jtensors (src/fgimic/jfield.f90:114-129).
PyTorch provides device memory, index arithmetic and streams.
"""
import ctypes
import math
import time

import numpy as np

from . import lib

ARCH = dict(frequencies=64, features=160, hidden=128, layers=4, outputs=9)


class Network:
    """Owner of one gimic_mlp handle.  Points and tensors are torch float64
    CUDA tensors; kernels run on torch's current stream."""

    def __init__(self, sigma, center_scale, out_scale, seed=0):
        import torch
        self._lib = lib.load()
        h = ctypes.c_void_p()
        cs = np.ascontiguousarray(center_scale, dtype=np.float64)
        os_ = np.ascontiguousarray(out_scale, dtype=np.float64)
        assert cs.size == 4 and os_.size == 9
        lib.check(self._lib.gimic_mlp_create(ctypes.byref(h), float(sigma), lib.ptr(cs),
                                             lib.ptr(os_), int(seed)))
        self._h = h
        self.sigma, self.center_scale, self.out_scale = float(sigma), cs, os_
        lib.check(self._lib.gimic_mlp_set_stream(
            self._h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def close(self):
        if self._h:
            lib.check(self._lib.gimic_mlp_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparams(self):
        n = ctypes.c_longlong()
        lib.check(self._lib.gimic_mlp_info(self._h, ctypes.byref(n), None))
        return n.value

    def freqs(self):
        f = np.zeros((ARCH["frequencies"], 3), dtype=np.float32)
        lib.check(self._lib.gimic_mlp_info(self._h, None, lib.ptr(f)))
        return f

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        lib.check(self._lib.gimic_mlp_params(self._h, lib.ptr(p), 0))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        lib.check(self._lib.gimic_mlp_params(self._h, lib.ptr(p), 1))

    @staticmethod
    def _dev(t, cols):
        import torch
        assert t.is_cuda and t.dtype == torch.float64 and t.shape[-1] == cols
        return t.contiguous()

    def loss_grad(self, r, target, eps):
        r, target = self._dev(r, 3), self._dev(target, 9)
        loss = ctypes.c_double()
        g = np.zeros(self.nparams, dtype=np.float32)
        lib.check(self._lib.gimic_mlp_loss_grad(self._h, r.shape[0], ctypes.c_void_p(r.data_ptr()),
                                                ctypes.c_void_p(target.data_ptr()), float(eps),
                                                ctypes.byref(loss), lib.ptr(g)))
        return loss.value, g

    def train_step(self, r, target, eps, lr, want_loss=False):
        r, target = self._dev(r, 3), self._dev(target, 9)
        loss = ctypes.c_double()
        lib.check(self._lib.gimic_mlp_train_step(
            self._h, r.shape[0], ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(target.data_ptr()),
            float(eps), float(lr), ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def predict(self, r, out=None, unfused=False):
        import torch
        r = self._dev(r, 3)
        if out is None:
            out = torch.empty((r.shape[0], 9), dtype=torch.float64, device=r.device)
        if unfused:
            lib.check(self._lib.gimic_mlp_predict_unfused(self._h, r.shape[0],
                                                          ctypes.c_void_p(r.data_ptr()),
                                                          ctypes.c_void_p(out.data_ptr())))
        else:
            lib.check(self._lib.gimic_mlp_predict(self._h, r.shape[0], ctypes.c_void_p(r.data_ptr()),
                                                  ctypes.c_void_p(out.data_ptr()), None))
        return out


def rel_cell_error(pred, exact, cell, ncells):
    """per-cell sqrt(sum |dJ|^2 / sum |J|^2) and held-out counts (torch)"""
    import torch
    d2 = ((pred - exact) ** 2).sum(dim=1)
    j2 = (exact ** 2).sum(dim=1)
    num = torch.zeros(ncells, dtype=torch.float64, device=pred.device).index_add_(0, cell, d2)
    den = torch.zeros(ncells, dtype=torch.float64, device=pred.device).index_add_(0, cell, j2)
    cnt = torch.zeros(ncells, dtype=torch.int64, device=pred.device).index_add_(
        0, cell, torch.ones_like(cell))
    err = torch.where(den > 0, torch.sqrt(num / den.clamp_min(1e-300)),
                      torch.full_like(den, float("inf")))
    return err, cnt


class Surrogate:
    """Network + exact-kernel fallback over an axis-aligned box."""

    def __init__(self, ctx, coords, lo, hi, spin="total", exclusion=0.4, tol=1e-3, ncell=8,
                 min_holdout=8, sigma=3.0, seed=0, precision="mixed"):
        import torch
        self.ctx, self.spin, self.precision = ctx, spin, precision
        self.dev = torch.device("cuda")
        self.coords = torch.as_tensor(np.asarray(coords, dtype=np.float64).reshape(-1, 3),
                                      device=self.dev)
        self.lo = np.asarray(lo, dtype=np.float64)
        self.hi = np.asarray(hi, dtype=np.float64)
        self.exclusion, self.tol, self.ncell, self.min_holdout = exclusion, tol, ncell, min_holdout
        self.sigma, self.seed = sigma, seed
        self.net = None
        self.cell_ok = None
        self.stats = {}

    # -- geometry -----------------------------------------------------------
    def cell_index(self, r):
        import torch
        lo = torch.as_tensor(self.lo, device=r.device)
        ext = torch.as_tensor(self.hi - self.lo, device=r.device)
        ijk = ((r - lo) / ext * self.ncell).floor().long().clamp_(0, self.ncell - 1)
        return ijk[:, 0] + self.ncell * (ijk[:, 1] + self.ncell * ijk[:, 2])

    def near_nucleus(self, r, chunk=1 << 20):
        import torch
        out = torch.empty(r.shape[0], dtype=torch.bool, device=r.device)
        for s in range(0, r.shape[0], chunk):
            d = torch.cdist(r[s:s + chunk], self.coords)
            out[s:s + chunk] = (d <= self.exclusion).any(dim=1)
        return out

    def sample(self, n, near_frac=0.5, near_sigma=1.0, seed=0):
        """n points in the box: (1 - near_frac) uniform, near_frac Gaussian
        clouds (std near_sigma bohr) around the nuclei, which is where the
        density and the current are large; points inside the exclusion
        spheres are redrawn (the exact kernel owns them)"""
        import torch
        g = torch.Generator(device=self.dev).manual_seed(int(seed))
        lo = torch.as_tensor(self.lo, device=self.dev)
        ext = torch.as_tensor(self.hi - self.lo, device=self.dev)
        out = torch.empty((0, 3), dtype=torch.float64, device=self.dev)
        while out.shape[0] < n:
            m = int((n - out.shape[0]) * 1.2) + 64
            nn = int(m * near_frac)
            u = lo + ext * torch.rand((m - nn, 3), generator=g, dtype=torch.float64, device=self.dev)
            a = torch.randint(0, self.coords.shape[0], (nn,), generator=g, device=self.dev)
            v = self.coords[a] + near_sigma * torch.randn((nn, 3), generator=g,
                                                          dtype=torch.float64, device=self.dev)
            p = torch.cat([u, v])
            keep = ((p >= lo) & (p <= lo + ext)).all(dim=1) & ~self.near_nucleus(p)
            out = torch.cat([out, p[keep]])
        return out[torch.randperm(out.shape[0], generator=g, device=self.dev)[:n]]

    # -- training -----------------------------------------------------------
    def fit(self, n_train=2_000_000, steps=4000, batch=16384, lr=2e-3, lr_min=2e-5, eps_frac=1e-3,
            holdout=0.01, near_frac=0.5, near_sigma=1.0, log_every=0):
        """sample, label with the exact kernel, train, validate per cell"""
        import torch
        t0 = time.perf_counter()
        pts = self.sample(n_train, near_frac=near_frac, near_sigma=near_sigma, seed=self.seed)
        J = self.ctx.jtensor_device(pts, spin=self.spin, precision=self.precision)
        torch.cuda.synchronize()
        t_label = time.perf_counter() - t0
        nh = max(1, int(round(holdout * n_train)))
        r_tr, j_tr = pts[nh:], J[nh:]
        r_ho, j_ho = pts[:nh], J[:nh]
        # network scales: centre/half-extent of the box, per-component RMS
        ctr = 0.5 * (self.lo + self.hi)
        scale = 0.5 * float((self.hi - self.lo).max())
        rms = torch.sqrt((j_tr ** 2).mean(dim=0)).cpu().numpy()
        out_scale = np.where(rms > 0, rms, 1.0)
        self.net = Network(self.sigma, np.r_[ctr, scale], out_scale, seed=self.seed)
        tn = torch.sqrt(((j_tr / torch.as_tensor(out_scale, device=self.dev)) ** 2).sum(dim=1))
        eps = eps_frac * float(tn.mean())
        g = torch.Generator(device=self.dev).manual_seed(int(self.seed) + 1)
        ntr = r_tr.shape[0]
        hist = []
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        perm = torch.randperm(ntr, generator=g, device=self.dev)
        pos = 0
        for it in range(steps):
            if pos + batch > ntr:
                perm = torch.randperm(ntr, generator=g, device=self.dev)
                pos = 0
            idx = perm[pos:pos + batch]
            pos += batch
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / steps))
            want = log_every and (it % log_every == 0 or it == steps - 1)
            loss = self.net.train_step(r_tr[idx], j_tr[idx], eps, lr_t, want_loss=bool(want))
            if want:
                hist.append((it, loss))
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        pred = self.net.predict(r_ho)
        cells = self.cell_index(r_ho)
        nc = self.ncell ** 3
        err, cnt = rel_cell_error(pred, j_ho, cells, nc)
        self._cell_err, self._cell_cnt = err, cnt
        self.cell_ok = (err <= self.tol) & (cnt >= self.min_holdout)
        tot = float(torch.sqrt(((pred - j_ho) ** 2).sum() / (j_ho ** 2).sum()))
        self.stats = dict(n_train=int(ntr), n_holdout=int(nh), steps=steps, batch=batch,
                          label_s=t_label, train_s=t_train, eps=eps, sigma=self.sigma,
                          holdout_rel_err=tot, cells=nc,
                          cells_served=int(self.cell_ok.sum()), loss_history=hist,
                          params=self.net.nparams)
        return self.stats

    # -- evaluation ---------------------------------------------------------
    def set_tolerance(self, tol):
        """re-select the served cells for another held-out tolerance"""
        self.tol = tol
        self.cell_ok = (self._cell_err <= tol) & (self._cell_cnt >= self.min_holdout)
        self.stats["cells_served"] = int(self.cell_ok.sum())

    def inside(self, r):
        """points in the closed training box (cell_index clamps the others
        into the edge cells, where the network was never checked)"""
        import torch
        lo = torch.as_tensor(self.lo, device=r.device)
        hi = torch.as_tensor(self.hi, device=r.device)
        return ((r >= lo) & (r <= hi)).all(dim=1)

    def served_mask(self, r):
        """the network serves a point only inside the box, in a cell whose
        held-out error met the tolerance, and beyond the exclusion radius"""
        return self.cell_ok[self.cell_index(r)] & self.inside(r) & ~self.near_nucleus(r)

    def evaluate(self, r, out=None):
        """J at every point of r: network where served, exact elsewhere;
        returns (J, served mask)"""
        import torch
        if self.net is None:
            raise RuntimeError("surrogate not trained")
        r = r.contiguous()
        if out is None:
            out = torch.empty((r.shape[0], 9), dtype=torch.float64, device=r.device)
        if not bool(self.cell_ok.any()):
            # no cell validated: the exact kernel everywhere
            self.ctx.jtensor_device(r, spin=self.spin, precision=self.precision, out=out)
            return out, torch.zeros(r.shape[0], dtype=torch.bool, device=r.device)
        served = self.served_mask(r)
        if bool(served.all()):
            self.net.predict(r, out=out)
            return out, served
        # the network on the served points only, the exact kernel on the rest
        sv = served.nonzero().squeeze(1)
        if sv.numel():
            out[sv] = self.net.predict(r[sv].contiguous())
        ex = (~served).nonzero().squeeze(1)
        out[ex] = self.ctx.jtensor_device(r[ex].contiguous(), spin=self.spin,
                                          precision=self.precision)
        return out, served

    def audit(self, r, J, served, nsample=100_000, seed=7):
        """exact-kernel check of a random sample of the served points:
        relative error sqrt(sum |dJ|^2 / sum |J|^2) and the worst cell"""
        import torch
        idx = served.nonzero().squeeze(1)
        if idx.numel() == 0:
            return dict(sampled=0, rel_err=0.0, worst_cell=0.0)
        g = torch.Generator(device=r.device).manual_seed(seed)
        if idx.numel() > nsample:
            idx = idx[torch.randperm(idx.numel(), generator=g, device=r.device)[:nsample]]
        ex = self.ctx.jtensor_device(r[idx].contiguous(), spin=self.spin, precision="fp64")
        d = J[idx] - ex
        rel = float(torch.sqrt((d ** 2).sum() / (ex ** 2).sum()))
        err, cnt = rel_cell_error(J[idx], ex, self.cell_index(r[idx]), self.ncell ** 3)
        worst = float(err[cnt >= self.min_holdout].max()) if bool((cnt >= self.min_holdout).any()) \
            else 0.0
        return dict(sampled=int(idx.numel()), rel_err=rel, worst_cell=worst)
