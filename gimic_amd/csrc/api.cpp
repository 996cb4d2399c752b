// C-ABI entry points: the reference's drop-in gimic_* functions on an implicit
// global context, and the batched handle API (include/gimic_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gimic_hip.h"
#include "device_common.h"
#include "host_basis.h"
#include "quadrature.h"

using namespace gimic;

namespace gimic {
// message of the last failing C-ABI call on this thread (shared by all entry points)
std::string &error_slot() {
    static thread_local std::string s;
    return s;
}
}  // namespace gimic

namespace {

std::string &g_error = gimic::error_slot();

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? GIMIC_ENOMEM : GIMIC_EDEVICE,
                    std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f) {
    try {
        f();
        g_error.clear();
        return GIMIC_OK;
    } catch (const Error &e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_error = "host allocation failed";
        return GIMIC_ENOMEM;
    } catch (const std::exception &e) {
        g_error = e.what();
        return GIMIC_EIO;
    }
}

// Device buffer: an owned allocation, or (view) a range of memory owned
// elsewhere, such as the device scratch of the allocation-free entry point.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool own = true;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && own) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        own = true;
    }
    void view(void *q, size_t n) {
        release();
        p = q;
        bytes = n;
        own = false;
    }
    void alloc(size_t n, hipStream_t s, bool zero = true) {
        if (n <= bytes && p) {
            if (zero) hip_check(hipMemsetAsync(p, 0, n, s), "hipMemsetAsync");
            return;
        }
        release();
        hip_check(hipMalloc(&p, std::max<size_t>(n, 256)), "hipMalloc");
        bytes = n;
        if (zero) hip_check(hipMemsetAsync(p, 0, n, s), "hipMemsetAsync");
    }
    template <class T>
    T *as() const { return (T *)p; }
};

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Device memory for the basis values and gradients of one chunk of points
// (one pass of basis evaluation, contraction and assembly).  The default is
// a fixed 8 GiB, modest beside other users of the device; the large setting
// (gimic_ctx_set_chunk_bytes(ctx, -1)) is a third of the device's memory,
// at most kChunkBytesMax.  The total, not the free memory, sets the large
// one, so the chunks (and with them the results' last bits, which the
// tiles' block lists touch) are the same on every run on one device model.
constexpr size_t kChunkBytesDefault = (size_t)8 << 30;
constexpr size_t kChunkBytesMax = (size_t)96 << 30;

// Bump allocator over a device range (the scratch of the allocation-free
// entry point): carve() hands out 256-byte aligned views and clears them
// with a kernel when asked (no hipMemsetAsync, so the calls may sit in a
// captured graph); nothing is freed until reset().
struct Arena {
    unsigned char *base = nullptr;
    size_t size = 0, used = 0;
    void reset(size_t to = 0) { used = to; }
    void carve(DevBuf &b, size_t bytes, bool zero, hipStream_t st);
};

// The work buffers of one evaluation: the context's own (grown as needed and
// kept), or views carved from an Arena.
struct Workspace {
    DevBuf phi64, grad64, phis, grad32, part;
    // spatial tiles of a chunk: sorted points, their caller indices, sort
    // scratch, and the screening block list of every contraction row tile
    DevBuf spts, perm, bpart, sortscr, bcount, blist, toff, k8count, k8list, bzero;
    // FP64 pass near nuclei: selection index list, their points, selection
    // scratch and count
    DevBuf near_idx, near_pts, near_scr, near_cnt;
};

void Arena::carve(DevBuf &b, size_t bytes, bool zero, hipStream_t st) {
    if (!(b.p && !b.own && b.bytes >= bytes)) {
        const size_t off = round_up(used, 256), n = round_up(std::max<size_t>(bytes, 1), 256);
        if (off > size || n > size - off)
            throw Error(GIMIC_ENOMEM, "device scratch exhausted (" + std::to_string(size >> 20) +
                                          " MiB)");
        b.view(base + off, bytes);
        used = off + n;
    }
    if (zero && bytes > 0)
        hip_check(launch_fill_zero((double *)b.p, round_up(bytes, 8) / 8, st), "fill_zero");
}

// Expected cost of a function order: the mean over sample points of
// (number of live 8-function groups)^2, the live (K-group, N-group) pairs
// of a row tile at the point, which is what the contraction kernels walk.
// live[p * nsh + s]: shell s reaches sample p.
template <class NOut>
double order_cost(const Basis &b, const std::vector<int> &order, const std::vector<char> &live,
                  int npt, NOut nout_of) {
    const int nsh = (int)b.shells.size();
    std::vector<int> fpos(nsh), fcnt(nsh);
    int pos = 0;
    for (int si : order) {
        fpos[si] = pos;
        fcnt[si] = nout_of(b.shells[si]);
        pos += fcnt[si];
    }
    const int ngrp = (pos + 7) / 8;
    std::vector<int> stamp(ngrp, -1);
    double sum = 0.0;
    for (int p = 0; p < npt; p++) {
        const char *lv = live.data() + (size_t)p * nsh;
        int n = 0;
        for (int s = 0; s < nsh; s++) {
            if (!lv[s]) continue;
            for (int k = fpos[s] >> 3; k <= (fpos[s] + fcnt[s] - 1) >> 3; k++)
                if (stamp[k] != p) {
                    stamp[k] = p;
                    n++;
                }
        }
        sum += (double)n * n;
    }
    return sum / npt;
}

// Candidate orders: the plain order by decreasing radius, and geometric
// radius buckets of ratio q and phase ph (bucket floor(log r / log q + ph),
// largest first; within a bucket by atom rank, then decreasing radius).
// Scored by order_cost on deterministic sample points spread over the
// molecule's bounding box widened by 1 bohr, with every radius widened by
// 0.5 bohr for the extent of a row tile; the first best candidate wins.
template <class NOut>
std::vector<int> choose_function_order(const Basis &b, const std::vector<int> &arank,
                                       NOut nout_of) {
    const int nsh = (int)b.shells.size(), nat = b.natoms();
    std::vector<int> base(nsh);
    for (int i = 0; i < nsh; i++) base[i] = i;
    std::stable_sort(base.begin(), base.end(), [&](int x, int y) {
        const Shell &p = b.shells[x], &q = b.shells[y];
        if (p.radius != q.radius) return p.radius > q.radius;
        return arank[p.atom] < arank[q.atom];
    });
    if (nsh < 2 || nat < 2) return base;
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = hi[d] = b.coords[d];
        for (int a = 1; a < nat; a++) {
            lo[d] = std::min(lo[d], b.coords[3 * a + d]);
            hi[d] = std::max(hi[d], b.coords[3 * a + d]);
        }
        lo[d] -= 1.0;
        hi[d] += 1.0;
    }
    // about 4096 samples, fewer for very large bases (live holds npt * nsh bytes)
    const int npt = (int)std::max<long long>(256, std::min<long long>(4096, 24000000LL / nsh));
    std::vector<char> live((size_t)npt * nsh);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto unif = [&st] {  // xorshift64*, fixed seed: the choice repeats run to run
        st ^= st >> 12;
        st ^= st << 25;
        st ^= st >> 27;
        return (double)((st * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int p = 0; p < npt; p++) {
        double x[3];
        for (int d = 0; d < 3; d++) x[d] = lo[d] + unif() * (hi[d] - lo[d]);
        for (int s = 0; s < nsh; s++) {
            const Shell &sh = b.shells[s];
            double r2 = 0.0;
            for (int d = 0; d < 3; d++) {
                const double e = x[d] - b.coords[3 * sh.atom + d];
                r2 += e * e;
            }
            const double r = sh.radius + 0.5;
            live[(size_t)p * nsh + s] = r2 <= r * r;
        }
    }
    std::vector<int> best = base;
    double best_cost = order_cost(b, base, live, npt, nout_of);
    const double qs[] = {1.25, 1.4, 1.6, 1.8, 2.0, 2.5, 3.2, 4.5};
    const double phs[] = {0.0, 0.25, 0.5, 0.75};
    std::vector<double> key(nsh);
    for (double q : qs)
        for (double ph : phs) {
            for (int i = 0; i < nsh; i++)
                key[i] = std::floor(std::log(b.shells[i].radius) / std::log(q) + ph);
            std::vector<int> o(nsh);
            for (int i = 0; i < nsh; i++) o[i] = i;
            std::stable_sort(o.begin(), o.end(), [&](int x, int y) {
                if (key[x] != key[y]) return key[x] > key[y];
                const Shell &p = b.shells[x], &r = b.shells[y];
                if (arank[p.atom] != arank[r.atom]) return arank[p.atom] < arank[r.atom];
                return p.radius > r.radius;
            });
            const double c = order_cost(b, o, live, npt, nout_of);
            if (c < best_cost) {
                best_cost = c;
                best = o;
            }
        }
    return best;
}

}  // namespace

struct gimic_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // host copies of pipelined batches
    size_t chunk_bytes = 0;  // basis buffers of one chunk of points
    size_t chunk_bytes_large = 0;  // the large setting: a third of the device memory
    Basis basis;
    double screen_thrs = 1e-6;
    int giao = 1, diamag = 1, paramag = 1;
    int nset = 0;
    std::vector<double> dens;  // nset*4*n*n column-major, as given
    int lda = 0, ncol_pad = 0;
    DevBuf shells, exps, ncc, coords, rf64, rf32;
    // right operands per (spin list, precision)
    struct Mats {
        DevBuf b64;
        DevBuf bs3;  // BF16 parts in LDS-image blocks (pre-split mixed kernel)
        int nmatset = 0;
    };
    std::map<std::pair<std::vector<int>, int>, std::unique_ptr<Mats>> mats;
    // evaluation workspace, and host-array staging of the batch calls
    Workspace ws;
    DevBuf pts_tmp, jt_tmp, contrib, planes, pparams, offs, outs;
    // internal function order (see setup_device): fun_orig[i] is the caller's
    // index of internal function i; per 32-function block the shells with a
    // function in it (blk_beg ranges into blk_sh)
    std::vector<int> fun_orig;
    DevBuf blk_beg, blk_sh;
    // property quadrature (gimic_ctx_properties): points + weights, tensors,
    // partition offsets, nuclei, sums
    DevBuf prop_in, prop_jt, prop_offs, prop_nuc, prop_out;
    // plane-grid descriptors (gimic_flux_plane_grids)
    DevBuf gspec;
    // mixed precision: points within fp64_radius bohr of a nucleus are
    // evaluated again in FP64 (0: off); points refined by the last mixed call
    double fp64_radius = 0.4;
    long long refined_last = 0;
    // optional per-kernel timing with HIP events on the launch stream:
    // [0] basis evaluation, [1] contraction, [2] finalize
    bool timing = false;
    // mixed-precision contraction kernel chosen by timing (-1: not yet tuned;
    // pre-split operands with direct-to-LDS staging: 3 and 4 with 4
    // wavefronts and 128 rows, double or triple buffered, 5 with 4
    // wavefronts and 256 rows, 8 and 9 with 8 wavefronts and 256 rows,
    // double or triple buffered) and the timings (basis evaluation +
    // contraction) that chose it.  Every variant, and the FP64 kernel
    // (variant 0 here), walks the screening block lists
    int mixed_variant = -1;
    int mixed_split = 0;  // N-split count of the pre-split kernel (0: by workgroup count)
    // block-list statistics (gimic_ctx_set_tile_stats): tiles, sum n, sum n^2, rows per tile
    bool tile_stats = false;
    double tile_acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    float tune_ms[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    std::vector<std::pair<int, float>> tune_log;  // (variant * 100 + split, ms)
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[3];
    double ms[3] = {0.0, 0.0, 0.0};
    long long launches[3] = {0, 0, 0};
    // flux planes as one captured hipGraph launch, reused while its key
    // (sizes, flags, kernel choice and every device address it bakes in)
    // matches
    bool graphs = false;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uintptr_t> graph_key;
    long long graph_builds = 0, graph_launches = 0;

    ~gimic_ctx() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (auto &v : ev)
            for (auto &e : v) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (copy_stream) {
            (void)hipStreamSynchronize(copy_stream);
            (void)hipStreamDestroy(copy_stream);
        }
    }
};

namespace {

void setup_device(gimic_ctx *c) {
    hip_check(hipGetDevice(&c->device), "hipGetDevice");
    hip_check(hipStreamCreate(&c->stream), "hipStreamCreate");
    size_t mem_free = 0, mem_total = 0;
    hip_check(hipMemGetInfo(&mem_free, &mem_total), "hipMemGetInfo");
    c->chunk_bytes_large =
        std::max<size_t>((size_t)1 << 30, std::min(kChunkBytesMax, mem_total / 3));
    c->chunk_bytes = kChunkBytesDefault;
    Basis &b = c->basis;
    finalize_basis(b, c->screen_thrs);
    c->lda = (int)round_up((size_t)b.nbf, kKPad);
    c->ncol_pad = (int)round_up((size_t)b.nbf, 32);
    // Internal function order: shells grouped into classes of like screening
    // radius (geometric buckets, largest first), within a class by atom
    // along a nearest-neighbour path through the molecule, then by
    // decreasing radius; so a 32-function block holds functions of like
    // range centred close together and screening zeroes whole blocks over
    // a compact tile of points.  The bucket ratio and phase are chosen per
    // basis by the expected count of live block pairs (order_cost).  Within
    // a shell the order is the caller's.  J is a sum over function pairs,
    // so only the summation order changes.
    const int nsh = (int)b.shells.size(), nat = b.natoms();
    std::vector<int> arank(std::max(nat, 1), 0);
    {
        std::vector<char> used(nat, 0);
        int cur = 0;
        for (int k = 0; k < nat; k++) {
            arank[cur] = k;
            used[cur] = 1;
            int best = -1;
            double bd = 0.0;
            for (int j = 0; j < nat; j++) {
                if (used[j]) continue;
                double d2 = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double e = b.coords[3 * j + d] - b.coords[3 * cur + d];
                    d2 += e * e;
                }
                if (best < 0 || d2 < bd) {
                    best = j;
                    bd = d2;
                }
            }
            if (best < 0) break;
            cur = best;
        }
    }
    auto nout_of = [&](const Shell &s) { return b.spherical ? 2 * s.l + 1 : ncart(s.l); };
    const std::vector<int> order = choose_function_order(b, arank, nout_of);
    std::vector<int> fnew(nsh, 0);
    c->fun_orig.assign(b.nbf, 0);
    {
        int pos = 0;
        for (int si : order) {
            const Shell &s = b.shells[si];
            fnew[si] = pos;
            for (int f = 0; f < nout_of(s); f++) c->fun_orig[pos + f] = s.fout + f;
            pos += nout_of(s);
        }
    }
    // Device shell array: ordered by l, then by output position.  The basis
    // kernels hand consecutive shells of a point to consecutive lanes, so a
    // wavefront then runs one l branch (one polynomial and transform) at a
    // time instead of all of them with most lanes masked off.
    std::vector<int> dorder(nsh), dpos(nsh);
    for (int i = 0; i < nsh; i++) dorder[i] = i;
    std::stable_sort(dorder.begin(), dorder.end(), [&](int x, int y) {
        if (b.shells[x].l != b.shells[y].l) return b.shells[x].l < b.shells[y].l;
        return fnew[x] < fnew[y];
    });
    for (int k = 0; k < nsh; k++) dpos[dorder[k]] = k;
    const int nblk = c->lda / 32;
    std::vector<int> blk_beg(nblk + 1, 0), blk_sh;  // entries index the device array
    for (int k = 0; k < nblk; k++) {
        blk_beg[k] = (int)blk_sh.size();
        for (int si : order) {
            const int f0 = fnew[si], f1 = f0 + nout_of(b.shells[si]);
            if (f0 < 32 * (k + 1) && f1 > 32 * k) blk_sh.push_back(dpos[si]);
        }
    }
    blk_beg[nblk] = (int)blk_sh.size();
    if (blk_sh.empty()) blk_sh.push_back(0);
    std::vector<DevShell> sh;
    for (int k = 0; k < nsh; k++) {
        const Shell &s = b.shells[dorder[k]];
        sh.push_back(DevShell{s.atom, s.l, s.nprim, s.poff, fnew[dorder[k]], nout_of(s), s.radius});
    }
    hipStream_t st = c->stream;
    auto up = [&](DevBuf &d, const void *src, size_t n) {
        d.alloc(n, st, false);
        hip_check(hipMemcpyAsync(d.p, src, n, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    };
    up(c->shells, sh.data(), sh.size() * sizeof(DevShell));
    up(c->exps, b.exps.data(), b.exps.size() * sizeof(double));
    up(c->ncc, b.ncc.data(), b.ncc.size() * sizeof(double));
    up(c->coords, b.coords.data(), b.coords.size() * sizeof(double));
    up(c->blk_beg, blk_beg.data(), blk_beg.size() * sizeof(int));
    up(c->blk_sh, blk_sh.data(), blk_sh.size() * sizeof(int));
    std::vector<double> rf(3 * (size_t)c->ncol_pad, 0.0);
    std::vector<float> rf32(3 * (size_t)c->ncol_pad, 0.f);
    for (int i = 0; i < nsh; i++) {
        const Shell &s = b.shells[i];
        for (int f = 0; f < nout_of(s); f++)
            for (int d = 0; d < 3; d++) {
                rf[(size_t)d * c->ncol_pad + fnew[i] + f] = b.coords[3 * s.atom + d];
                rf32[(size_t)d * c->ncol_pad + fnew[i] + f] = (float)b.coords[3 * s.atom + d];
            }
    }
    up(c->rf64, rf.data(), rf.size() * sizeof(double));
    up(c->rf32, rf32.data(), rf32.size() * sizeof(float));
    std::vector<double> sph;
    for (int l = 2; l <= kDevMaxL; l++) {
        auto T = sph_transform(l);
        sph.insert(sph.end(), T.begin(), T.end());
    }
    hip_check(upload_basis_tables(sph.data(), (int)sph.size(), st), "upload_basis_tables");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// FP64 density family of one spin case: 4 n*n column-major blocks.
std::vector<double> spin_family(const gimic_ctx *c, int spin) {
    const size_t n = (size_t)c->basis.nbf, blk = 4 * n * n;
    if (c->nset == 0) throw Error(GIMIC_ESTATE, "no density set");
    if (c->nset == 1) {
        if (spin == GIMIC_SPIN_BETA)
            throw Error(GIMIC_ESTATE, "beta current requested, but not open-shell system!");
        if (spin == GIMIC_SPIN_SPINDENS)
            throw Error(GIMIC_ESTATE, "spindens requested, but not open-shell system!");
        return std::vector<double>(c->dens.begin(), c->dens.begin() + blk);
    }
    std::vector<double> out(blk);
    const double *a = c->dens.data(), *bb = c->dens.data() + blk;
    for (size_t i = 0; i < blk; i++) {
        switch (spin) {
            case GIMIC_SPIN_ALPHA: out[i] = a[i]; break;
            case GIMIC_SPIN_BETA: out[i] = bb[i]; break;
            case GIMIC_SPIN_SPINDENS: out[i] = a[i] - bb[i]; break;
            default: out[i] = a[i] + bb[i]; break;
        }
    }
    return out;
}

gimic_ctx::Mats &get_mats(gimic_ctx *c, const std::vector<int> &spins, int prec) {
    auto key = std::make_pair(spins, prec);
    auto it = c->mats.find(key);
    if (it != c->mats.end()) return *it->second;
    const size_t n = (size_t)c->basis.nbf, ld = (size_t)c->lda, nc = (size_t)c->ncol_pad;
    const size_t matsz = nc * ld, setsz = kNMat * matsz;
    auto m = std::make_unique<gimic_ctx::Mats>();
    m->nmatset = (int)spins.size();
    std::vector<double> B(setsz * spins.size(), 0.0);
    // atom of each caller-ordered function; the internal order is fun_orig
    std::vector<int> atom_of(n);
    for (auto &s : c->basis.shells) {
        const int nout = c->basis.spherical ? 2 * s.l + 1 : ncart(s.l);
        for (int f = 0; f < nout; f++) atom_of[s.fout + f] = s.atom;
    }
    const std::vector<int> &fo = c->fun_orig;
    for (size_t si = 0; si < spins.size(); si++) {
        std::vector<double> fam = spin_family(c, spins[si]);
        double *Bs = &B[si * setsz];
        // matrix j column nu (K = mu contiguous), internal order:
        // D(mu,nu) = fam[blk + fo[mu] + n*fo[nu]].
        // The GIAO right operands are W_d(mu,nu) = D(mu,nu) (R_nu - R_mu)_d:
        // phi W_d equals R_d (phi D) - phi diag(R_d) D without the
        // cancellation between those two terms (and is zero on one-atom blocks).
        for (size_t nu = 0; nu < n; nu++) {
            const size_t onu = (size_t)fo[nu];
            const double *Rn = &c->basis.coords[3 * atom_of[onu]];
            for (size_t mu = 0; mu < n; mu++) {
                const size_t omu = (size_t)fo[mu];
                const double D = fam[omu + n * onu];
                const double *Rm = &c->basis.coords[3 * atom_of[omu]];
                Bs[0 * matsz + nu * ld + mu] = D;
                for (int b = 1; b < 4; b++)
                    Bs[b * matsz + nu * ld + mu] = fam[(size_t)b * n * n + omu + n * onu];
                for (int d = 0; d < 3; d++)
                    Bs[(4 + d) * matsz + nu * ld + mu] = D * (Rn[d] - Rm[d]);
            }
        }
    }
    hipStream_t st = c->stream;
    if (prec == GIMIC_PREC_FP64) {
        m->b64.alloc(B.size() * sizeof(double), st, false);
        hip_check(hipMemcpyAsync(m->b64.p, B.data(), B.size() * sizeof(double),
                                 hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
    } else {
        std::vector<float> b32(B.begin(), B.end());
        // the FP32 values as their first kBParts BF16 parts (three parts:
        // x = h + m + l exactly), one pair block per (set, 8-function K-group
        // gk, 8-function N-group gn), [part][matrix][n][k]; rows of K-groups
        // of ngn + 1 blocks (the last zero) and a zero row last: the mixed
        // kernels gather the 16 blocks of an N-tile's K-step into their LDS
        // buffers
        const size_t nkg = ld / 8, ngn = nc / 8, pe = kMx2PairElems;
        const size_t setel = mx2_set_elems((int)ld, (int)nc);
        std::vector<uint16_t> bs(spins.size() * setel, 0);
        auto bf = [](float x) {
            uint32_t u;
            std::memcpy(&u, &x, 4);
            u += 0x7FFFu + ((u >> 16) & 1u);
            return (uint16_t)(u >> 16);
        };
        auto fl = [](uint16_t h) {
            const uint32_t u = (uint32_t)h << 16;
            float x;
            std::memcpy(&x, &u, 4);
            return x;
        };
        for (size_t si = 0; si < spins.size(); si++)
            for (size_t mat = 0; mat < (size_t)kNMat; mat++)
                for (size_t nu = 0; nu < nc; nu++) {
                    const float *src = &b32[si * setsz + mat * matsz + nu * ld];
                    const size_t gn = nu / 8, nn = nu % 8;
                    for (size_t mu = 0; mu < ld; mu++) {
                        const size_t gk = mu / 8, kk = mu % 8;
                        uint16_t *dst = &bs[si * setel + (gk * (ngn + 1) + gn) * pe];
                        float x = src[mu];
                        for (int p = 0; p < kBParts; p++) {
                            const uint16_t h = bf(x);
                            dst[((p * kNMat + mat) * 8 + nn) * 8 + kk] = h;
                            x -= fl(h);
                        }
                    }
                }
        m->bs3.alloc(bs.size() * 2, st, false);
        hip_check(hipMemcpyAsync(m->bs3.p, bs.data(), bs.size() * 2, hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    auto &ref = *m;
    c->mats[key] = std::move(m);
    return ref;
}

// Run `launch` between two events on stream st when timing is on.
template <class F>
void timed(gimic_ctx *c, int which, hipStream_t st, F &&launch) {
    if (!c->timing) {
        launch();
        return;
    }
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    c->ev[which].push_back({a, b});
    hip_check(hipEventRecord(a, st), "hipEventRecord");
    launch();
    hip_check(hipEventRecord(b, st), "hipEventRecord");
}

struct FluxDev {
    const double *w = nullptr;
    const int *plane = nullptr;
    const double *pnorm = nullptr, *pbfield = nullptr, *pcenter = nullptr, *pradius = nullptr;
    double *contrib = nullptr;
    double *contrib_mod = nullptr;
    double *contrib_acid = nullptr;
};

enum class EvalMode {
    Eager,    // allocate as needed, tune, launch
    Reserve,  // allocate every buffer the launches need and tune; launch nothing else
    Capture,  // kernel launches only (no allocation, memset, timing or sync)
    Scratch,  // work buffers carved from an Arena: no allocation, tuning, timing or
              // statistics (the allocation-free entry point)
};

// Where an evaluation's work buffers come from: the context's workspace
// (ws null), another Workspace, and in EvalMode::Scratch views carved from
// `arena`, at most max_rows rows per chunk.
struct EvalWork {
    Workspace *ws = nullptr;
    Arena *arena = nullptr;
    size_t max_rows = 0;
};

// The mixed kernel and N-split count of an untuned context in Scratch mode:
// 8 wavefronts, triple buffered, four N-ranges (the tuner's usual choice on
// large grids).  Scratch mode splits a row tile over at most
// kScratchMaxSplit N-ranges, which bounds the partial sums it plans for.
constexpr int kDefaultVariant = 9, kDefaultSplit = 4;
constexpr int kScratchMaxSplit = 32;

// Row tile of the mixed kernels: 128 rows (4 wavefronts x 32), 256 (4 x 64
// or 8 x 32)
int mixed_tile_rows(int variant) { return variant == 5 || variant >= 8 ? 256 : 128; }

// Row tile of the FP64 kernel: 128 rows, or 64 for chunks of up to
// kSmallF64Points points (the FP64 pass over the points next to nuclei, the
// nuclei's share of property grids): there neighbouring points are far
// apart, and the smaller tiles list a fifth fewer group pairs on the
// 60-atom cage.  The screening lists are sized for the smallest tile.
constexpr int kMinTileRows = 64;
constexpr size_t kSmallF64Points = 131072;
int f64_tile_rows(size_t points) { return points <= kSmallF64Points ? 64 : kWgRows; }

// Arena bytes one Scratch-mode chunk of `rows` rows carves (every carve
// rounded to 256 bytes), with `split` N-ranges per row tile
size_t scratch_eval_bytes(const gimic_ctx *c, int prec, int nmatset, size_t rows, int split) {
    const size_t ld = (size_t)c->lda, nblk = ld / 32, t = rows / kMinTileRows;
    auto r = [](size_t b) { return round_up(std::max<size_t>(b, 1), 256); };
    size_t n = prec == GIMIC_PREC_FP64 ? r(rows * ld * 8) + r(3 * rows * ld * 8)
                                       : r(rows * ld * 2 * kAParts) + r(3 * rows * ld * 4);
    n += r(rows * 24) + r(rows * 4) + r(256 * 6 * 8) + r((size_t)sort_scratch_ints((int)rows) * 4);
    n += r(t * 4) + r(t * nblk * 4) + r((t + 1) * 4) + r(t * 4) + r(t * nblk * 16) + r(12 * 1024);
    n += r((size_t)nmatset * split * kNQ * rows * 8);
    return n;
}

// Largest chunk (a multiple of the precision's row granularity) whose
// Scratch-mode buffers fit in `bytes`; 0 if not even one fits
size_t scratch_rows(const gimic_ctx *c, int prec, int nmatset, size_t bytes, int split) {
    const size_t gran = prec == GIMIC_PREC_FP64 ? (size_t)kWgRows : 256;
    size_t lo = 0, hi = gran;
    while (scratch_eval_bytes(c, prec, nmatset, hi, split) <= bytes) {
        lo = hi;
        hi *= 2;
    }
    while (hi - lo > gran) {
        const size_t mid = round_up((lo + hi) / 2, gran);
        if (mid >= hi) break;
        if (scratch_eval_bytes(c, prec, nmatset, mid, split) <= bytes) lo = mid;
        else hi = mid;
    }
    return lo;
}

void refine_near_nuclei(gimic_ctx *c, int64_t npts, const double *d_r, double *d_jt,
                        const std::vector<int> &spins, hipStream_t st, const FluxDev *flux,
                        EvalWork work = {});

// Evaluate J for npts device points into d_jt (nspin blocks of 9*npts).
// With out_idx (a subset run), point i's tensor and flux contribution go to
// index out_idx[i] of output arrays sized for out_npts points.  A mixed
// evaluation in eager mode ends with the FP64 pass over the points next to
// a nucleus (refine_near_nuclei); a captured or Scratch-mode one leaves it
// to the caller.
void evaluate(gimic_ctx *c, int64_t npts, const double *d_r, double *d_jt,
              const std::vector<int> &spins, int prec, hipStream_t st, const FluxDev *flux,
              EvalMode mode = EvalMode::Eager, const long long *out_idx = nullptr,
              int64_t out_npts = 0, EvalWork work = {}) {
    if (npts <= 0) return;
    if (prec != GIMIC_PREC_FP64 && prec != GIMIC_PREC_MIXED)
        throw Error(GIMIC_EINVAL, "unknown precision");
    for (int s : spins)
        if (s < 0 || s > 3) throw Error(GIMIC_EINVAL, "invalid spin case");
    const bool capture = mode == EvalMode::Capture;
    const bool scratch = mode == EvalMode::Scratch;
    if ((capture || scratch) && !c->mats.count(std::make_pair(spins, prec)))
        throw Error(GIMIC_ESTATE, capture ? "capture before reserve: right operands missing"
                                          : "right operands of this spin case and precision "
                                            "are not resident (one allocating call builds them)");
    if (scratch && (!work.arena || work.max_rows == 0))
        throw Error(GIMIC_EINVAL, "scratch evaluation without a planned chunk");
    Workspace &W = work.ws ? *work.ws : c->ws;
    // an eager mixed evaluation builds the FP64 operands of its FP64 pass
    // near nuclei up front, so that a later allocation-free call finds both
    if (prec == GIMIC_PREC_MIXED && mode == EvalMode::Eager && !out_idx && c->fp64_radius > 0.0)
        (void)get_mats(c, spins, GIMIC_PREC_FP64);
    auto &M = get_mats(c, spins, prec);
    // per-kernel event timing only for eager evaluation
    struct TimingOff {
        gimic_ctx *c;
        bool saved;
        ~TimingOff() { c->timing = saved; }
    } timing_guard{c, c->timing};
    if (mode != EvalMode::Eager) c->timing = false;
    auto grow = [&](DevBuf &b, size_t bytes, bool zero) {
        if (scratch) {
            work.arena->carve(b, bytes, zero, st);
            return;
        }
        if (capture) {
            if (b.bytes < bytes || !b.p) throw Error(GIMIC_ESTATE, "capture before reserve");
            return;
        }
        b.alloc(bytes, st, zero);
    };
    // mixed kernel: the tuned choice, or in Scratch mode without one the default
    int mv = c->mixed_variant, msplit = c->mixed_split;
    if (scratch && prec == GIMIC_PREC_MIXED && mv < 0) {
        mv = kDefaultVariant;
        msplit = kDefaultSplit;
    }

    const size_t ld = (size_t)c->lda;
    // chunk of points per pass: basis values + gradients of one chunk stay
    // within c->chunk_bytes (FP64: 32 bytes per point and function; mixed:
    // split values 2 kAParts + FP32 gradients 12 bytes)
    const size_t per_row = ld * (prec == GIMIC_PREC_FP64 ? 32 : 2 * kAParts + 12);
    // row granularity: a multiple of every contraction kernel's workgroup
    // height (128; 256 for the 256-row pre-split kernels)
    const size_t gran = prec == GIMIC_PREC_FP64 ? (size_t)kWgRows : 256;
    const size_t rows_all = round_up((size_t)npts, gran);
    size_t chunk = scratch ? work.max_rows
                           : std::max<size_t>(gran, (c->chunk_bytes / per_row) / gran * gran);
    // as few chunks as the cap allows, all of one size (a short last chunk
    // would be a launch of its own with a tail that fills the GPU badly)
    auto even_chunks = [&](size_t cap) {
        const size_t nmin = (rows_all + cap - 1) / cap;
        return round_up((rows_all + nmin - 1) / nmin, gran);
    };
    chunk = even_chunks(chunk);
    // padding columns and rows of the basis buffers must read as zero: they
    // are cleared when (re)allocated (every Scratch-mode call carves them
    // afresh) and never written afterwards.  On an allocation failure in
    // eager mode the other precision's basis buffers are released first, then
    // the chunk is halved (smaller chunks change only the tiles' block lists,
    // i.e. the results' last bits)
    for (;;) {
        const size_t rc = chunk;
        try {
            if (prec == GIMIC_PREC_FP64) {
                grow(W.phi64, rc * ld * 8, W.phi64.bytes < rc * ld * 8);
                grow(W.grad64, 3 * rc * ld * 8, W.grad64.bytes < 3 * rc * ld * 8);
            } else {
                grow(W.phis, rc * ld * 2 * kAParts, W.phis.bytes < rc * ld * 2 * kAParts);
                grow(W.grad32, 3 * rc * ld * 4, W.grad32.bytes < 3 * rc * ld * 4);
            }
            break;
        } catch (const Error &e) {
            if (e.code != GIMIC_ENOMEM || mode != EvalMode::Eager) throw;
            (void)hipGetLastError();  // the failed hipMalloc is not a launch error
            if (prec == GIMIC_PREC_FP64 && (W.phis.p || W.grad32.p)) {
                W.phis.release();
                W.grad32.release();
            } else if (prec != GIMIC_PREC_FP64 && (W.phi64.p || W.grad64.p)) {
                W.phi64.release();
                W.grad64.release();
            } else if (chunk > gran) {
                chunk = even_chunks(std::max(gran, chunk / 2 / gran * gran));
            } else {
                throw;
            }
        }
    }
    const size_t rows_cap = chunk;
    const long long gplane = (long long)(rows_cap * ld);
    // spatial tiles: sorted points, caller indices, sort scratch, and block
    // lists for row tiles of kMinTileRows rows or more
    const int nblk = c->lda / 32;
    grow(W.spts, rows_cap * 3 * 8, false);
    grow(W.perm, rows_cap * 4, false);
    grow(W.bpart, 256 * 6 * 8, false);
    grow(W.sortscr, (size_t)sort_scratch_ints((int)rows_cap) * 4, false);
    grow(W.bcount, rows_cap / kMinTileRows * 4, false);
    grow(W.blist, rows_cap / kMinTileRows * nblk * 4, false);
    grow(W.toff, (rows_cap / kMinTileRows + 1) * 4, false);
    grow(W.k8count, rows_cap / kMinTileRows * 4, false);
    grow(W.k8list, rows_cap / kMinTileRows * nblk * 4 * 4, false);
    grow(W.bzero, 12 * 1024, W.bzero.bytes < 12 * 1024);  // zeroed once, never written
    const int nmatset = M.nmatset;
    // chunks in rotated order, the middle one first: the autotuner times the
    // first chunk it sees, and the middle of a grid around a molecule is
    // where the work is densest (each chunk's results are its own, so the
    // order changes nothing else)
    const size_t nchunks = ((size_t)npts + chunk - 1) / chunk;
    for (size_t ci = 0; ci < nchunks; ci++) {
        const size_t start = ((ci + nchunks / 2) % nchunks) * chunk;
        const int n_c = (int)std::min(chunk, (size_t)npts - start);
        const int rows = (int)round_up((size_t)n_c, gran);

        const double *spts = W.spts.as<double>();
        EvalArgs ea{};
        ea.pts = spts;
        ea.npts = n_c;
        ea.rows = rows;
        ea.gplane = gplane;
        ea.shells = c->shells.as<DevShell>();
        ea.nshell = (int)c->basis.shells.size();
        ea.exps = c->exps.as<double>();
        ea.ncc = c->ncc.as<double>();
        ea.coords = c->coords.as<double>();
        ea.turbomole = c->basis.turbomole;
        ea.spherical = c->basis.spherical;
        ea.lda = c->lda;
        ea.phi64 = W.phi64.as<double>();
        ea.grad64 = W.grad64.as<double>();
        ea.grad32 = W.grad32.as<float>();
        ea.phis = W.phis.as<__bf16>();
        ea.lmax = 0;
        for (auto &sh : c->basis.shells) ea.lmax = std::max(ea.lmax, sh.l);

        ContractArgs ca{};
        ca.rows = rows;
        ca.lda = c->lda;
        ca.ncol_pad = c->ncol_pad;
        ca.gplane = gplane;
        ca.nmatset = nmatset;
        ca.pts = spts;
        ca.npts = n_c;
        ca.a64 = W.phi64.as<double>();
        ca.g64 = W.grad64.as<double>();
        ca.b64 = M.b64.as<double>();
        ca.g32 = W.grad32.as<float>();
        ca.as3 = W.phis.as<__bf16>();
        ca.bs3 = M.bs3.as<__bf16>();
        ca.rf64 = c->rf64.as<double>();
        ca.rf32 = c->rf32.as<float>();
        ca.nblk = nblk;
        ca.k8count = W.k8count.as<int>();
        ca.k8list = W.k8list.as<int>();
        ca.bzero = W.bzero.as<__bf16>();
        // screening block lists for row tiles of tile_rows rows
        auto lists = [&](int tile_rows) {
            TileArgs ta{};
            ta.spts = spts;
            ta.npts = n_c;
            ta.rows = rows;
            ta.tile_rows = tile_rows;
            ta.nblk = nblk;
            ta.blk_beg = c->blk_beg.as<int>();
            ta.blk_sh = c->blk_sh.as<int>();
            ta.shells = c->shells.as<DevShell>();
            ta.coords = c->coords.as<double>();
            ta.bcount = W.bcount.as<int>();
            ta.blist = W.blist.as<int>();
            ta.k8count = W.k8count.as<int>();
            ta.k8list = W.k8list.as<int>();
            hip_check(launch_block_lists(ta, st), "block_lists");
            if (c->tile_stats && mode == EvalMode::Eager) {
                const int nt = rows / tile_rows;
                std::vector<int> cnt(nt), cnt8(nt);
                hip_check(hipMemcpyAsync(cnt.data(), ta.bcount, nt * sizeof(int),
                                         hipMemcpyDeviceToHost, st),
                          "hipMemcpyAsync");
                hip_check(hipMemcpyAsync(cnt8.data(), ta.k8count, nt * sizeof(int),
                                         hipMemcpyDeviceToHost, st),
                          "hipMemcpyAsync");
                hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
                for (int t = 0; t < nt; t++) {
                    const int k = cnt[t];
                    c->tile_acc[0] += 1.0;
                    c->tile_acc[1] += k;
                    c->tile_acc[2] += (double)k * k;
                    c->tile_acc[4] += (double)k * cnt8[t] / 4.0;
                    c->tile_acc[5] += (double)cnt8[t] * cnt8[t];
                }
                c->tile_acc[3] = tile_rows;
            }
        };
        // work split over N so a launch has at least ~8192 workgroups (32
        // per CU; small FP64 launches, such as the pass over the points next
        // to nuclei, ran 4.5% faster than at 2048 and 22% faster than at
        // 512), or the split count the tuner chose for the pre-split kernel
        auto plan = [&](int wg_rows, int tile_cols, int want_split = 0) {
            ca.ntiles = c->ncol_pad / tile_cols;
            const long long per = (long long)(rows / wg_rows) * nmatset;
            int nsplit = (int)std::min<long long>(ca.ntiles,
                                                  std::max<long long>(1, (8192 + per - 1) / per));
            if (want_split > 0) nsplit = std::min(want_split, ca.ntiles);
            if (scratch) nsplit = std::min(nsplit, kScratchMaxSplit);
            ca.tiles_per_split = (ca.ntiles + nsplit - 1) / nsplit;
            ca.nsplit = (ca.ntiles + ca.tiles_per_split - 1) / ca.tiles_per_split;
            grow(W.part, (size_t)nmatset * ca.nsplit * kNQ * rows * 8, false);
            ca.part = W.part.as<double>();
        };
        auto plan_variant = [&](int variant, int split = -1) {
            if (split < 0) split = msplit;
            if (variant == 0) plan(f64_tile_rows(n_c), kNT);
            else plan(mixed_tile_rows(variant), kMx2Cols / kNMat, split);
        };
        // mixed kernels: row-major partial sums (written whole by the launch),
        // and basis rows compressed to each row tile's live groups (lists
        // and offsets built before the basis evaluation)
        auto run = [&](int variant, int split = -1) {
            plan_variant(variant, split);
            if (variant == 0) lists(f64_tile_rows(n_c));  // the FP64 kernel walks lists over dense rows
            ca.toff = variant != 0 ? W.toff.as<int>() : nullptr;
            if (variant == 0) hip_check(launch_contract_f64(ca, f64_tile_rows(n_c), st), "contract_f64");
            else if (variant == 3) hip_check(launch_contract_mx2(ca, 2, 2, 4, st), "contract_mx2");
            else if (variant == 4) hip_check(launch_contract_mx2(ca, 3, 2, 4, st), "contract_mx2");
            else if (variant == 5) hip_check(launch_contract_mx2(ca, 2, 4, 4, st), "contract_mx2");
            else if (variant == 8) hip_check(launch_contract_mx2(ca, 2, 2, 8, st), "contract_mx2");
            else hip_check(launch_contract_mx2(ca, 3, 2, 8, st), "contract_mx2");
        };
        // basis evaluation: FP64 rows (mode 0), or the compressed split rows
        // the mixed kernels read (mode 2)
        auto eval = [&](int variant) {
            ea.toff = nullptr;
            if (variant != 0) {
                const int tr = mixed_tile_rows(variant);
                lists(tr);
                hip_check(launch_tile_offsets(W.k8count.as<int>(), rows / tr, W.toff.as<int>(), st),
                          "tile_offsets");
                ea.k8count = W.k8count.as<int>();
                ea.k8list = W.k8list.as<int>();
                ea.toff = W.toff.as<int>();
                ea.nblk = nblk;
                ea.tile_rows = tr;
            }
            hip_check(launch_basis_eval(ea, variant == 0 ? 0 : 2, st), "basis_eval");
        };
        // tensor assembly (and flux contributions) of this chunk from the
        // partial sums of the last contraction launch
        auto finalize = [&](int variant) {
            FinalizeArgs fa{};
            fa.part = W.part.as<double>();
            fa.nsplit = ca.nsplit;
            fa.part_rows = variant != 0 ? 1 : 0;
            fa.rows = rows;
            fa.npts = n_c;
            fa.nmatset = nmatset;
            fa.pts = spts;
            fa.perm = W.perm.as<int>();
            fa.giao = c->giao;
            fa.diamag = c->diamag;
            fa.paramag = c->paramag;
            // outputs: this chunk's slice, or (subset run) the whole arrays
            // addressed through out_idx
            const size_t o0 = out_idx ? 0 : start;
            fa.pidx = out_idx ? out_idx + start : nullptr;
            fa.jt = d_jt + 9 * o0;
            fa.jt_stride = 9 * (long long)(out_idx ? out_npts : npts);
            if (flux) {
                fa.w = flux->w + o0;
                fa.plane = flux->plane ? flux->plane + o0 : nullptr;
                fa.pnorm = flux->pnorm;
                fa.pbfield = flux->pbfield;
                fa.pcenter = flux->pcenter;
                fa.pradius = flux->pradius;
                fa.contrib = flux->contrib + o0;
                fa.contrib_mod = flux->contrib_mod ? flux->contrib_mod + o0 : nullptr;
                fa.contrib_acid = flux->contrib_acid ? flux->contrib_acid + o0 : nullptr;
            }
            hip_check(launch_finalize(fa, st), "finalize");
        };
        const bool tune = prec == GIMIC_PREC_MIXED && mv < 0;
        if (capture && tune) throw Error(GIMIC_ESTATE, "mixed kernel not tuned before capture");
        if (mode == EvalMode::Reserve && !tune) {
            plan_variant(prec == GIMIC_PREC_FP64 ? 0 : mv);
            continue;
        }
        // rows in Morton order of their cells (spts, perm)
        timed(c, 0, st, [&] {
            hip_check(launch_sort_points(d_r + 3 * start, n_c, W.spts.as<double>(),
                                         W.perm.as<int>(), W.bpart.as<double>(),
                                         W.sortscr.as<int>(), st),
                      "sort_points");
        });
        if (tune) {
            // autotune once per context on this chunk: basis evaluation,
            // contraction and assembly for the mixed kernels, each with
            // several N-split counts (more splits give more workgroups that
            // share each row tile's A through L2, fewer give each workgroup
            // more N-tiles to amortise its start, and fewer partial sums to
            // assemble); every candidate is timed three times and its
            // fastest run counts.  The fastest candidate replaces the
            // default only when it is faster by more than 1%: near-ties are
            // noise, and a chance pick of a near-tie ran 4% slower in the
            // steady state than the default on the 60-atom cage.
            const int cand[][2] = {{3, 4}, {4, 4}, {5, 4}, {8, 2}, {8, 4}, {8, 8},
                                   {9, 2}, {9, 4}, {9, 8}};
            c->tune_log.clear();
            float best = 0.f, def_ms = 0.f;
            int best_v = -1, best_split = 0;
            for (auto &cv : cand) {
                const int v = cv[0], sp = cv[1];
                hipEvent_t e0, e1;
                hip_check(hipEventCreate(&e0), "hipEventCreate");
                hip_check(hipEventCreate(&e1), "hipEventCreate");
                plan_variant(v, sp);  // allocate before the timed region
                float ms = 0.f;
                for (int rep = 0; rep < 3; rep++) {
                    hip_check(hipEventRecord(e0, st), "hipEventRecord");
                    eval(v);
                    run(v, sp);
                    finalize(v);
                    hip_check(hipEventRecord(e1, st), "hipEventRecord");
                    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float t = 0.f;
                    hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
                    ms = rep == 0 ? t : std::min(ms, t);
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                c->tune_log.push_back({v * 100 + sp, ms});
                if (c->tune_ms[v - 1] == 0.f || ms < c->tune_ms[v - 1]) c->tune_ms[v - 1] = ms;
                if (v == kDefaultVariant && sp == kDefaultSplit) def_ms = ms;
                if (best_v < 0 || ms < best) {
                    best = ms;
                    best_v = v;
                    best_split = sp;
                }
            }
            if (def_ms > 0.f && def_ms <= 1.01f * best) {
                best_v = kDefaultVariant;
                best_split = kDefaultSplit;
            }
            c->mixed_variant = best_v;
            c->mixed_split = best_split;
            mv = c->mixed_variant;
            msplit = c->mixed_split;
            if (mode == EvalMode::Reserve) {
                plan_variant(mv);
                continue;
            }
        }
        const int variant = prec == GIMIC_PREC_FP64 ? 0 : mv;
        timed(c, 0, st, [&] { eval(variant); });
        timed(c, 1, st, [&] { run(variant); });
        timed(c, 2, st, [&] { finalize(variant); });
    }
    if (prec == GIMIC_PREC_MIXED && mode == EvalMode::Eager && !out_idx)
        refine_near_nuclei(c, npts, d_r, d_jt, spins, st, flux);
}

// The FP64 pass of a mixed evaluation: the points within fp64_radius of a
// nucleus, where J is large and FP32 products lose the most relative
// accuracy, are selected on the device (in index order), evaluated again in
// FP64, and their tensors and flux contributions overwrite the mixed ones in
// place.  Synchronises once, to size the subset.  With an arena in `work`,
// its buffers and those of the FP64 evaluation are carved from the arena's
// free part.
void refine_near_nuclei(gimic_ctx *c, int64_t npts, const double *d_r, double *d_jt,
                        const std::vector<int> &spins, hipStream_t st, const FluxDev *flux,
                        EvalWork work) {
    c->refined_last = 0;
    if (!(c->fp64_radius > 0.0) || npts <= 0) return;
    if (npts > 0x7fffffffLL) throw Error(GIMIC_EINVAL, "too many points for the FP64 pass");
    Workspace &W = work.ws ? *work.ws : c->ws;
    auto get = [&](DevBuf &b, size_t bytes) {
        if (work.arena) work.arena->carve(b, bytes, false, st);
        else b.alloc(bytes, st, false);
    };
    get(W.near_idx, (size_t)npts * 8);
    get(W.near_scr, near_select_scratch(npts) * 4);
    get(W.near_cnt, 8);
    hip_check(launch_near_select(d_r, npts, c->coords.as<double>(), c->basis.natoms(),
                                 c->fp64_radius, W.near_idx.as<long long>(),
                                 W.near_cnt.as<long long>(), W.near_scr.as<int>(), st),
              "near_select");
    long long k = 0;
    hip_check(hipMemcpyAsync(&k, W.near_cnt.p, 8, hipMemcpyDeviceToHost, st), "copy count");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (k <= 0) return;
    get(W.near_pts, (size_t)k * 24);
    hip_check(launch_gather_points(d_r, W.near_idx.as<long long>(), k, W.near_pts.as<double>(), st),
              "gather_points");
    if (work.arena) {
        // the FP64 evaluation carves its own views after the selection's
        Workspace sub;
        Arena &A = *work.arena;
        const size_t used = round_up(A.used, 256);
        const size_t rows = scratch_rows(c, GIMIC_PREC_FP64, (int)spins.size(),
                                         A.size > used ? A.size - used : 0, kScratchMaxSplit);
        if (rows == 0) throw Error(GIMIC_ENOMEM, "device scratch too small for the FP64 pass");
        evaluate(c, k, W.near_pts.as<double>(), d_jt, spins, GIMIC_PREC_FP64, st, flux,
                 EvalMode::Scratch, W.near_idx.as<long long>(), npts, EvalWork{&sub, &A, rows});
    } else {
        evaluate(c, k, W.near_pts.as<double>(), d_jt, spins, GIMIC_PREC_FP64, st, flux,
                 EvalMode::Eager, W.near_idx.as<long long>(), npts);
    }
    c->refined_last = k;
}

// ---- host-array batches: copies beside the evaluation ----

// A host-array batch of more than kPipeMin points (gimic_jtensor_batch)
// runs in sub-batches: near-equal ones of at most kPipeHead points, then
// kPipeTail[0] and kPipeTail[1] points.  A helper thread makes the host
// copies on a second stream of default priority, in order: the points of
// sub-batch b + 2 go up and the tensors of sub-batch b come down while the
// caller's thread evaluates sub-batch b + 1 on the context's stream, so the
// PCIe copies (the host arrays are pageable: each copy holds the thread
// that issues it until it is done) overlap the kernels.  What stays exposed
// is the first sub-batch's points going up and the last one's tensors
// coming down, hence the short tail: the tensors of a sub-batch come down
// (about 7 ns a point into fresh pageable memory on MI355X's host link)
// while the next, a quarter its size, is evaluated (about 33 ns a point on
// the 60-atom cage), and each sub-batch costs about 1 ms of its own.  Each
// sub-batch has its own events: "up" (its points are on the device) is
// waited for by the context's stream, "done" (its tensors are written) by
// the copy stream.  A sub-batch is evaluated like a call of its own, so
// mixed results agree with an unsplit evaluation to the mixed tolerance and
// FP64 results to summation-order rounding.
constexpr int64_t kPipeHead = 1048576;
constexpr int64_t kPipeTail[2] = {262144, 65536};
constexpr int64_t kPipeMin = kPipeHead + kPipeTail[0] + kPipeTail[1];

void pipelined_batch(gimic_ctx *c, int64_t npts, const double *r, double *jt, int spin,
                     int precision) {
    hipStream_t st = c->stream;
    std::vector<int64_t> offs{0};  // sub-batch b: points offs[b] .. offs[b + 1]
    const int64_t head = npts - kPipeTail[0] - kPipeTail[1];
    const int64_t nhead = (head + kPipeHead - 1) / kPipeHead;
    for (int64_t k = 1; k <= nhead; k++) offs.push_back(head * k / nhead);
    offs.push_back(head + kPipeTail[0]);
    offs.push_back(npts);
    const int nb = (int)offs.size() - 1;
    auto first = [&](int b) { return offs[b]; };
    auto count = [&](int b) { return offs[b + 1] - offs[b]; };
    double *d_r = c->pts_tmp.as<double>(), *d_jt = c->jt_tmp.as<double>();
    if (!c->copy_stream) hip_check(hipStreamCreate(&c->copy_stream), "hipStreamCreate");
    hipStream_t cs = c->copy_stream;
    std::vector<hipEvent_t> up(nb, nullptr), done(nb, nullptr);
    struct Events {
        std::vector<hipEvent_t> *v[2];
        ~Events() {
            for (auto *w : v)
                for (hipEvent_t e : *w)
                    if (e) (void)hipEventDestroy(e);
        }
    } events{{&up, &done}};
    for (int b = 0; b < nb; b++) {
        hip_check(hipEventCreateWithFlags(&up[b], hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&done[b], hipEventDisableTiming), "hipEventCreate");
    }
    // copy jobs in the order the helper runs them: +b + 1 = points of b up,
    // -(b + 1) = tensors of b down (after its "done" event is recorded)
    std::mutex mu;
    std::condition_variable cv;
    std::vector<int> jobs;
    size_t next_job = 0;
    int ups = 0;          // sub-batches whose points are up and "up" recorded
    int dones = 0;        // sub-batches whose "done" the main thread recorded
    int downs = 0;        // sub-batches whose tensors are down
    bool stop = false;
    std::exception_ptr err;
    const int dev = c->device;
    std::thread helper([&] {
        try {
            hip_check(hipSetDevice(dev), "hipSetDevice");
            for (;;) {
                int job = 0;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || next_job < jobs.size(); });
                    if (next_job == jobs.size()) return;
                    job = jobs[next_job++];
                }
                if (job > 0) {
                    const int b = job - 1;
                    hip_check(hipMemcpyAsync(d_r + 3 * first(b), r + 3 * first(b),
                                             (size_t)count(b) * 24, hipMemcpyHostToDevice, cs),
                              "hipMemcpyAsync");
                    hip_check(hipEventRecord(up[b], cs), "hipEventRecord");
                    std::lock_guard<std::mutex> lk(mu);
                    ups = b + 1;
                } else {
                    const int b = -job - 1;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [&] { return stop || dones > b; });
                        if (dones <= b) return;
                    }
                    hip_check(hipStreamWaitEvent(cs, done[b], 0), "hipStreamWaitEvent");
                    hip_check(hipMemcpyAsync(jt + 9 * first(b), d_jt + 9 * first(b),
                                             (size_t)count(b) * 72, hipMemcpyDeviceToHost, cs),
                              "hipMemcpyAsync");
                    hip_check(hipStreamSynchronize(cs), "hipStreamSynchronize");
                    std::lock_guard<std::mutex> lk(mu);
                    downs = b + 1;
                }
                cv.notify_all();
            }
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            err = std::current_exception();
            stop = true;
            cv.notify_all();
        }
    });
    // ends the helper (on an error too) before the frame's state goes away
    struct Join {
        std::thread &t;
        std::mutex &m;
        std::condition_variable &v;
        bool &stop;
        ~Join() {
            {
                std::lock_guard<std::mutex> lk(m);
                stop = true;
            }
            v.notify_all();
            if (t.joinable()) t.join();
        }
    } join{helper, mu, cv, stop};
    auto push = [&](int job) {
        std::lock_guard<std::mutex> lk(mu);
        jobs.push_back(job);
        cv.notify_all();
    };
    push(1);
    push(2);
    long long refined = 0;  // points of the FP64 pass over all sub-batches
    for (int b = 0; b < nb; b++) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return err || ups > b; });
            if (err) break;
        }
        hip_check(hipStreamWaitEvent(st, up[b], 0), "hipStreamWaitEvent");
        evaluate(c, count(b), d_r + 3 * first(b), d_jt + 9 * first(b), {spin}, precision, st,
                 nullptr);
        refined += c->refined_last;
        hip_check(hipEventRecord(done[b], st), "hipEventRecord");
        {
            std::lock_guard<std::mutex> lk(mu);
            dones = b + 1;
        }
        push(-(b + 1));
        if (b + 2 < nb) push(b + 3);
    }
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return err || downs == nb; });
    }
    if (err) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamSynchronize(cs);
        std::rethrow_exception(err);
    }
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (precision == GIMIC_PREC_MIXED) c->refined_last = refined;
}

// ---- allocation-free evaluation (gimic_jtensor_batch_device_noalloc) ----

// Work memory of the allocation-free entry point: one copy on every GPU the
// library runs on, held until the process ends.  Host code uses it only
// through the device address hipGetSymbolAddress returns (scratch_address).
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current
unsigned char *scratch_address() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw Error(GIMIC_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw Error(GIMIC_EDEVICE, "no device address for the scratch array");
        g_scratch_addr[dev] = p;
    }
    return (unsigned char *)g_scratch_addr[dev];
}

// Makes `dev` current for the guard's lifetime
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        hip_check(hipGetDevice(&prev), "hipGetDevice");
        if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// ---- drop-in global state (gimic_interface.f90:21-24) ----
gimic_ctx *g_ctx = nullptr;
double g_magnet[3] = {0.0, 0.0, 0.0};
int g_spin = GIMIC_SPIN_TOTAL;
int g_uhf_next = 0;
double g_thrs_next = 1e-6;  // SCREEN_THRS, globals.f90:58

void report(const char *where) {
    if (!g_error.empty()) std::fprintf(stderr, "gimic: %s: %s\n", where, g_error.c_str());
}

}  // namespace

extern "C" {

const char *gimic_last_error(void) { return g_error.c_str(); }

int gimic_ctx_create_files(gimic_ctx **out, const char *molfile, const char *densfile, int uhf,
                           double screen_thrs) {
    return guarded([&] {
        if (!out || !molfile) throw Error(GIMIC_EINVAL, "null argument");
        auto c = std::make_unique<gimic_ctx>();
        c->basis = read_mol(molfile);
        c->screen_thrs = screen_thrs;
        if (densfile) {
            c->nset = uhf ? 2 : 1;
            c->dens = read_xdens(densfile, c->basis, c->nset, uhf != 0);
        }
        setup_device(c.get());
        *out = c.release();
    });
}

int gimic_ctx_create(gimic_ctx **out, int natoms, const double *coords, int nshell,
                     const int *shell_atom, const int *shell_l, const int *shell_nprim,
                     const int *shell_poff, const double *exps, const double *coefs,
                     int turbomole, int spherical, double screen_thrs) {
    return guarded([&] {
        if (!out || !coords || !shell_atom || !shell_l || !shell_nprim || !shell_poff || !exps ||
            !coefs)
            throw Error(GIMIC_EINVAL, "null argument");
        auto c = std::make_unique<gimic_ctx>();
        try {
            c->basis = make_basis(natoms, coords, nshell, shell_atom, shell_l, shell_nprim,
                                  shell_poff, exps, coefs, turbomole != 0, spherical != 0);
        } catch (const std::runtime_error &e) {
            throw Error(GIMIC_EINVAL, e.what());
        }
        c->screen_thrs = screen_thrs;
        setup_device(c.get());
        *out = c.release();
    });
}

int gimic_ctx_set_density(gimic_ctx *c, int nset, const double *dens) {
    return guarded([&] {
        if (!c || !dens || (nset != 1 && nset != 2)) throw Error(GIMIC_EINVAL, "bad argument");
        const size_t n = (size_t)c->basis.nbf;
        c->nset = nset;
        c->dens.assign(dens, dens + (size_t)nset * 4 * n * n);
        c->mats.clear();
    });
}

int gimic_ctx_set_flags(gimic_ctx *c, int giao, int diamag, int paramag) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        c->giao = giao != 0;
        c->diamag = diamag != 0;
        c->paramag = paramag != 0;
    });
}

int gimic_ctx_info(const gimic_ctx *c, int *nbf, int *nbf_cart, int *natoms, int *nshell,
                   int *nset) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        if (nbf) *nbf = c->basis.nbf;
        if (nbf_cart) *nbf_cart = c->basis.nbf_cart;
        if (natoms) *natoms = c->basis.natoms();
        if (nshell) *nshell = (int)c->basis.shells.size();
        if (nset) *nset = c->nset;
    });
}

int gimic_ctx_atoms(const gimic_ctx *c, double *coords, double *charges) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        if (coords) std::memcpy(coords, c->basis.coords.data(), c->basis.coords.size() * 8);
        if (charges) std::memcpy(charges, c->basis.charges.data(), c->basis.charges.size() * 8);
    });
}

int gimic_ctx_atom_symbols(const gimic_ctx *c, char *sym) {
    return guarded([&] {
        if (!c || !sym) throw Error(GIMIC_EINVAL, "null argument");
        for (int a = 0; a < c->basis.natoms(); a++) {
            const std::string &s = a < (int)c->basis.symbols.size() ? c->basis.symbols[a] : "X ";
            sym[2 * a] = s.size() > 0 ? s[0] : ' ';
            sym[2 * a + 1] = s.size() > 1 ? s[1] : ' ';
        }
    });
}

int gimic_ctx_set_timing(gimic_ctx *c, int on) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        c->timing = on != 0;
        for (int k = 0; k < 3; k++) {
            for (auto &e : c->ev[k]) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
            c->ev[k].clear();
            c->ms[k] = 0.0;
            c->launches[k] = 0;
        }
    });
}

int gimic_ctx_get_timing(gimic_ctx *c, double *ms, long long *launches) {
    return guarded([&] {
        if (!c || !ms || !launches) throw Error(GIMIC_EINVAL, "null argument");
        for (int k = 0; k < 3; k++) {
            for (auto &e : c->ev[k]) {
                hip_check(hipEventSynchronize(e.second), "hipEventSynchronize");
                float t = 0.f;
                hip_check(hipEventElapsedTime(&t, e.first, e.second), "hipEventElapsedTime");
                c->ms[k] += t;
                c->launches[k] += 1;
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
            c->ev[k].clear();
            ms[k] = c->ms[k];
            launches[k] = c->launches[k];
        }
    });
}

int gimic_ctx_tuning(const gimic_ctx *c, int *variant, int *split, double *ms) {
    return guarded([&] {
        if (!c || !variant || !split || !ms) throw Error(GIMIC_EINVAL, "null argument");
        *variant = c->mixed_variant;
        *split = c->mixed_split;
        for (int k = 0; k < 9; k++) ms[k] = c->tune_ms[k];
    });
}

int gimic_ctx_destroy(gimic_ctx *c) {
    return guarded([&] { delete c; });
}

int gimic_jtensor_multi_device(gimic_ctx *c, int64_t npts, const double *d_r, double *d_jt,
                               int nspin, const int *spins, int precision, void *stream) {
    return guarded([&] {
        if (!c || npts < 0 || (npts > 0 && (!d_r || !d_jt)) || nspin < 1 || !spins)
            throw Error(GIMIC_EINVAL, "bad argument");
        hipStream_t st = stream ? (hipStream_t)stream : c->stream;
        evaluate(c, npts, d_r, d_jt, std::vector<int>(spins, spins + nspin), precision, st,
                 nullptr);
    });
}

int gimic_jtensor_batch_device(gimic_ctx *c, int64_t npts, const double *d_r, double *d_jt,
                               int spin, int precision, void *stream) {
    return gimic_jtensor_multi_device(c, npts, d_r, d_jt, 1, &spin, precision, stream);
}

int gimic_jtensor_batch_device_noalloc(gimic_ctx *c, int64_t npts, const double *d_r,
                                       double *d_jt, int spin, int precision, void *stream) {
    return guarded([&] {
        if (!c || npts < 0 || (npts > 0 && (!d_r || !d_jt))) throw Error(GIMIC_EINVAL, "bad argument");
        if (spin < 0 || spin > 3) throw Error(GIMIC_EINVAL, "invalid spin case");
        if (precision != GIMIC_PREC_FP64 && precision != GIMIC_PREC_MIXED)
            throw Error(GIMIC_EINVAL, "unknown precision");
        if (npts == 0) return;
        DeviceGuard device(c->device);
        hipStream_t st = stream ? (hipStream_t)stream : c->stream;
        const std::vector<int> spins{spin};
        const bool refine = precision == GIMIC_PREC_MIXED && c->fp64_radius > 0.0;
        if (!c->mats.count(std::make_pair(spins, precision)) ||
            (refine && !c->mats.count(std::make_pair(spins, (int)GIMIC_PREC_FP64))))
            throw Error(GIMIC_ESTATE,
                        "right operands of this spin case and precision are not resident: one "
                        "gimic_jtensor_batch_device call with them (one point suffices) builds "
                        "them");
        Arena A;
        A.base = scratch_address();
        A.size = kScratchBytes;
        // chunks of points whose work buffers fill the scratch
        const int split = precision == GIMIC_PREC_FP64 ? kScratchMaxSplit
                          : c->mixed_variant < 0       ? kDefaultSplit
                          : c->mixed_split > 0         ? c->mixed_split
                                                       : kScratchMaxSplit;
        const size_t rows = scratch_rows(c, precision, 1, kScratchBytes, split);
        if (rows == 0)
            throw Error(GIMIC_ENOMEM, "the " + std::to_string(kScratchBytes >> 20) +
                                          " MiB device scratch holds less than one row tile at " +
                                          std::to_string(c->basis.nbf) + " basis functions");
        long long refined = 0;
        for (int64_t start = 0; start < npts; start += (int64_t)rows) {
            const int64_t n = std::min<int64_t>((int64_t)rows, npts - start);
            Workspace W;
            A.reset();
            evaluate(c, n, d_r + 3 * start, d_jt + 9 * start, spins, precision, st, nullptr,
                     EvalMode::Scratch, nullptr, 0, EvalWork{&W, &A, rows});
            if (refine) {
                // the mixed pass's buffers are free once its launches are queued
                Workspace W2;
                A.reset();
                refine_near_nuclei(c, n, d_r + 3 * start, d_jt + 9 * start, spins, st, nullptr,
                                   EvalWork{&W2, &A, 0});
                refined += c->refined_last;
            }
        }
        c->refined_last = refined;
    });
}

int gimic_jtensor_batch(gimic_ctx *c, int64_t npts, const double *r, double *jt, int spin,
                        int precision) {
    return guarded([&] {
        if (!c || npts < 0 || (npts > 0 && (!r || !jt))) throw Error(GIMIC_EINVAL, "bad argument");
        if (npts == 0) return;
        hipStream_t st = c->stream;
        c->pts_tmp.alloc((size_t)npts * 3 * 8, st, false);
        c->jt_tmp.alloc((size_t)npts * 9 * 8, st, false);
        if (npts > kPipeMin) {
            pipelined_batch(c, npts, r, jt, spin, precision);
            return;
        }
        hip_check(hipMemcpyAsync(c->pts_tmp.p, r, (size_t)npts * 24, hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
        evaluate(c, npts, c->pts_tmp.as<double>(), c->jt_tmp.as<double>(), {spin}, precision, st,
                 nullptr);
        hip_check(hipMemcpyAsync(jt, c->jt_tmp.p, (size_t)npts * 72, hipMemcpyDeviceToHost, st),
                  "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}

}  // extern "C"

namespace {

// Work buffers of a plane batch: points then weights (pts_tmp), tensors,
// per-point contributions (current, signed modulus, ACID: every point is
// written by finalize_kernel, so no clearing), plane index, plane
// parameters, offsets and per-plane sums.
void flux_buffers(gimic_ctx *c, int64_t npts, int nplanes) {
    hipStream_t st = c->stream;
    const size_t np = (size_t)std::max<int64_t>(npts, 1);
    c->pts_tmp.alloc(np * 24 + np * 8, st, false);
    c->jt_tmp.alloc(np * 72, st, false);
    c->contrib.alloc(np * 24, st, false);
    c->planes.alloc(np * 4, st, false);
    c->pparams.alloc((size_t)nplanes * 80, st, false);
    c->offs.alloc((size_t)(nplanes + 1) * 8, st, false);
    c->outs.alloc((size_t)nplanes * 72, st, false);
}

// Plane flux over the points, weights and plane indices that `prepare(st)`
// leaves in the work buffers (a device launch; in graph mode it is captured
// with the rest).  offs and the per-plane parameters are host arrays.
// `source` tells the graph cache which kind of input `prepare` reads.
template <class Prep>
void flux_run(gimic_ctx *c, int64_t npts, int nplanes, const int64_t *offs, const double *normals,
              const double *bfields, const double *centers, const double *radii, int spin,
              int precision, double *out, double *out_mod, double *out_acid, double *jt,
              std::vector<uintptr_t> source, Prep &&prepare) {
    hipStream_t st = c->stream;
    std::vector<double> params(10 * (size_t)nplanes);
    for (int k = 0; k < nplanes; k++) {
        for (int d = 0; d < 3; d++) {
            params[3 * k + d] = normals[3 * k + d];
            params[3 * nplanes + 3 * k + d] = bfields[3 * k + d];
            params[6 * nplanes + 3 * k + d] = centers[3 * k + d];
        }
        params[9 * nplanes + k] = radii[k];
    }
    const size_t np = (size_t)std::max<int64_t>(npts, 1);
    double *d_r = c->pts_tmp.as<double>(), *d_w = d_r + 3 * np;
    hip_check(hipMemcpyAsync(c->pparams.p, params.data(), params.size() * 8,
                             hipMemcpyHostToDevice, st), "copy params");
    hip_check(hipMemcpyAsync(c->offs.p, offs, (size_t)(nplanes + 1) * 8, hipMemcpyHostToDevice, st),
              "copy offs");
    FluxDev fd;
    fd.w = d_w;
    fd.plane = c->planes.as<int>();
    const double *pp = c->pparams.as<double>();
    fd.pnorm = pp;
    fd.pbfield = pp + 3 * nplanes;
    fd.pcenter = pp + 6 * nplanes;
    fd.pradius = pp + 9 * nplanes;
    fd.contrib = c->contrib.as<double>();
    fd.contrib_mod = out_mod ? c->contrib.as<double>() + np : nullptr;
    fd.contrib_acid = out_acid ? c->contrib.as<double>() + 2 * np : nullptr;
    auto reduce = [&] {
        for (int k = 0; k < 3; k++) {
            if ((k == 1 && !out_mod) || (k == 2 && !out_acid)) continue;
            hip_check(launch_plane_reduce(c->contrib.as<double>() + k * np,
                                          (const long long *)c->offs.p, nplanes,
                                          c->outs.as<double>() + 3 * k * nplanes, st),
                      "plane_reduce");
        }
    };
    // the FP64 pass over points next to a nucleus runs eagerly after a
    // captured mixed evaluation, and the reductions after it
    const bool refine = precision == GIMIC_PREC_MIXED && c->fp64_radius > 0.0;
    if (!c->graphs) {
        prepare(st);
        evaluate(c, npts, d_r, c->jt_tmp.as<double>(), {spin}, precision, st, &fd);
        reduce();
    } else {
        // the inputs' preparation, all planes, every chunk and the per-plane
        // reductions as one graph.  An untuned mixed kernel is tuned in the
        // Reserve pass on the points themselves, so they are prepared first
        if (precision == GIMIC_PREC_MIXED && c->mixed_variant < 0) prepare(st);
        evaluate(c, npts, d_r, c->jt_tmp.as<double>(), {spin}, precision, st, &fd,
                 EvalMode::Reserve);
        auto &M = get_mats(c, {spin}, precision);
        std::vector<uintptr_t> key = {
            (uintptr_t)npts, (uintptr_t)nplanes, (uintptr_t)precision, (uintptr_t)spin,
            (uintptr_t)(out_mod != nullptr), (uintptr_t)(out_acid != nullptr),
            (uintptr_t)c->giao, (uintptr_t)c->diamag, (uintptr_t)c->paramag,
            (uintptr_t)c->mixed_variant, (uintptr_t)st, (uintptr_t)c->pts_tmp.p,
            (uintptr_t)c->jt_tmp.p, (uintptr_t)c->contrib.p, (uintptr_t)c->planes.p,
            (uintptr_t)c->pparams.p, (uintptr_t)c->offs.p, (uintptr_t)c->outs.p,
            (uintptr_t)c->ws.phi64.p, (uintptr_t)c->ws.grad64.p, (uintptr_t)c->ws.grad32.p,
            (uintptr_t)c->ws.part.p, (uintptr_t)M.b64.p, (uintptr_t)c->ws.phis.p,
            (uintptr_t)M.bs3.p, (uintptr_t)c->mixed_split, (uintptr_t)c->ws.spts.p,
            (uintptr_t)c->ws.perm.p, (uintptr_t)c->ws.bpart.p, (uintptr_t)c->ws.sortscr.p,
            (uintptr_t)c->ws.bcount.p, (uintptr_t)c->ws.blist.p, (uintptr_t)c->ws.toff.p,
            (uintptr_t)c->ws.k8count.p, (uintptr_t)c->ws.k8list.p, (uintptr_t)c->ws.bzero.p,
            (uintptr_t)refine,
            (uintptr_t)c->chunk_bytes};
        key.insert(key.end(), source.begin(), source.end());
        if (!c->graph_exec || key != c->graph_key) {
            if (c->graph_exec) {
                hip_check(hipGraphExecDestroy(c->graph_exec), "hipGraphExecDestroy");
                c->graph_exec = nullptr;
            }
            hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
                      "hipStreamBeginCapture");
            hipGraph_t g = nullptr;
            try {
                prepare(st);
                evaluate(c, npts, d_r, c->jt_tmp.as<double>(), {spin}, precision, st, &fd,
                         EvalMode::Capture);
                if (!refine) reduce();
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
            hipError_t e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(e, "hipGraphInstantiate");
            c->graph_key = key;
            c->graph_builds++;
        }
        hip_check(hipGraphLaunch(c->graph_exec, st), "hipGraphLaunch");
        c->graph_launches++;
        if (refine) {
            refine_near_nuclei(c, npts, d_r, c->jt_tmp.as<double>(), {spin}, st, &fd);
            reduce();
        }
    }
    std::vector<double> res(9 * (size_t)nplanes);
    hip_check(hipMemcpyAsync(res.data(), c->outs.p, res.size() * 8, hipMemcpyDeviceToHost, st),
              "copy out");
    if (jt && npts > 0)
        hip_check(hipMemcpyAsync(jt, c->jt_tmp.p, (size_t)npts * 72, hipMemcpyDeviceToHost, st),
                  "copy jt");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    std::copy(res.begin(), res.begin() + 3 * nplanes, out);
    if (out_mod) std::copy(res.begin() + 3 * nplanes, res.begin() + 6 * nplanes, out_mod);
    if (out_acid)
        for (int k = 0; k < nplanes; k++) out_acid[k] = res[6 * (size_t)nplanes + 3 * k];
}

}  // namespace

extern "C" {

int gimic_flux_planes(gimic_ctx *c, int nplanes, const int64_t *offs, const double *r,
                      const double *w, const double *normals, const double *bfields,
                      const double *centers, const double *radii, int spin, int precision,
                      double *out, double *out_mod, double *out_acid, double *jt) {
    return guarded([&] {
        if (!c || nplanes < 1 || !offs || !r || !w || !normals || !bfields || !centers ||
            !radii || !out)
            throw Error(GIMIC_EINVAL, "null argument");
        if (offs[0] != 0) throw Error(GIMIC_EINVAL, "plane_offs[0] must be 0");
        for (int k = 0; k < nplanes; k++)
            if (offs[k + 1] < offs[k]) throw Error(GIMIC_EINVAL, "plane_offs must be non-decreasing");
        const int64_t npts = offs[nplanes];
        hipStream_t st = c->stream;
        flux_buffers(c, npts, nplanes);
        const size_t np = (size_t)std::max<int64_t>(npts, 1);
        double *d_r = c->pts_tmp.as<double>(), *d_w = d_r + 3 * np;
        if (npts > 0) {
            hip_check(hipMemcpyAsync(d_r, r, (size_t)npts * 24, hipMemcpyHostToDevice, st), "copy r");
            hip_check(hipMemcpyAsync(d_w, w, (size_t)npts * 8, hipMemcpyHostToDevice, st), "copy w");
        }
        flux_run(c, npts, nplanes, offs, normals, bfields, centers, radii, spin, precision, out,
                 out_mod, out_acid, jt, {(uintptr_t)0}, [&](hipStream_t s) {
                     hip_check(launch_plane_index((const long long *)c->offs.p, nplanes, npts,
                                                  c->planes.as<int>(), s),
                               "plane_index");
                 });
    });
}

int gimic_flux_plane_grids(gimic_ctx *c, int nplanes, const int *n1, const int *n2,
                           const double *origins, const double *axes, const double *p3w3,
                           const double *nodes1, const double *wts1, const double *nodes2,
                           const double *wts2, const double *normals, const double *bfields,
                           const double *centers, const double *radii, int spin, int precision,
                           double *out, double *out_mod, double *out_acid) {
    return guarded([&] {
        if (!c || nplanes < 1 || !n1 || !n2 || !origins || !axes || !p3w3 || !nodes1 || !wts1 ||
            !nodes2 || !wts2 || !normals || !bfields || !centers || !radii || !out)
            throw Error(GIMIC_EINVAL, "null argument");
        // host side: O(nplanes) offsets; the device generates the points
        std::vector<int64_t> poff(nplanes + 1, 0), off1(nplanes, 0), off2(nplanes, 0);
        int64_t s1 = 0, s2 = 0;
        for (int k = 0; k < nplanes; k++) {
            if (n1[k] < 1 || n2[k] < 1) throw Error(GIMIC_EINVAL, "every plane needs n1, n2 >= 1");
            off1[k] = s1;
            off2[k] = s2;
            s1 += n1[k];
            s2 += n2[k];
            poff[k + 1] = poff[k] + (int64_t)n1[k] * n2[k];
        }
        const int64_t npts = poff[nplanes];
        hipStream_t st = c->stream;
        flux_buffers(c, npts, nplanes);
        // one spec buffer: n1 | off1 off2 poff | origin axes p3w3 nodes1 wts1 nodes2 wts2
        const size_t ni = (size_t)nplanes, nl = 3 * ni + 1,
                     nd = 3 * ni + 9 * ni + 2 * ni + 2 * (size_t)s1 + 2 * (size_t)s2;
        std::vector<unsigned char> spec(round_up(ni * 4, 8) + nl * 8 + nd * 8);
        unsigned char *q = spec.data();
        std::memcpy(q, n1, ni * 4);
        int64_t *lq = (int64_t *)(q + round_up(ni * 4, 8));
        std::memcpy(lq, off1.data(), ni * 8);
        std::memcpy(lq + ni, off2.data(), ni * 8);
        std::memcpy(lq + 2 * ni, poff.data(), (ni + 1) * 8);
        double *dq = (double *)(lq + nl);
        size_t o = 0;
        auto put = [&](const double *src, size_t n) {
            std::memcpy(dq + o, src, n * 8);
            o += n;
        };
        put(origins, 3 * ni);
        put(axes, 9 * ni);
        put(p3w3, 2 * ni);
        put(nodes1, (size_t)s1);
        put(wts1, (size_t)s1);
        put(nodes2, (size_t)s2);
        put(wts2, (size_t)s2);
        c->gspec.alloc(spec.size(), st, false);
        hip_check(hipMemcpyAsync(c->gspec.p, spec.data(), spec.size(), hipMemcpyHostToDevice, st),
                  "copy grid spec");
        const unsigned char *dbase = (const unsigned char *)c->gspec.p;
        const long long *dl = (const long long *)(dbase + round_up(ni * 4, 8));
        const double *dd = (const double *)(dl + nl);
        PlaneGridArgs ga{};
        ga.nplanes = nplanes;
        ga.npts = npts;
        ga.n1 = (const int *)dbase;
        ga.off1 = dl;
        ga.off2 = dl + ni;
        ga.poff = dl + 2 * ni;
        ga.origin = dd;
        ga.axes = dd + 3 * ni;
        ga.p3w3 = dd + 12 * ni;
        ga.nodes1 = dd + 14 * ni;
        ga.wts1 = ga.nodes1 + s1;
        ga.nodes2 = ga.wts1 + s1;
        ga.wts2 = ga.nodes2 + s2;
        const size_t np = (size_t)std::max<int64_t>(npts, 1);
        double *d_r = c->pts_tmp.as<double>(), *d_w = d_r + 3 * np;
        flux_run(c, npts, nplanes, poff.data(), normals, bfields, centers, radii, spin, precision,
                 out, out_mod, out_acid, nullptr,
                 {(uintptr_t)1, (uintptr_t)c->gspec.p, (uintptr_t)s1, (uintptr_t)s2},
                 [&](hipStream_t s) {
                     hip_check(launch_plane_grid(ga, d_r, d_w, c->planes.as<int>(), s),
                               "plane_grid");
                 });
    });
}

int gimic_ctx_tuning_log(const gimic_ctx *c, int *n, int *codes, double *ms) {
    return guarded([&] {
        if (!c || !n || *n < 0 || (*n > 0 && (!codes || !ms))) throw Error(GIMIC_EINVAL, "bad argument");
        const int m = std::min<int>(*n, (int)c->tune_log.size());
        for (int k = 0; k < m; k++) {
            codes[k] = c->tune_log[k].first;
            ms[k] = c->tune_log[k].second;
        }
        *n = (int)c->tune_log.size();
    });
}

int gimic_ctx_set_mixed_kernel(gimic_ctx *c, int variant, int split) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        if (variant != 0 && variant != 3 && variant != 4 && variant != 5 && variant != 8 &&
            variant != 9)
            throw Error(GIMIC_EINVAL, "mixed kernel must be 0 (tune), 3, 4, 5, 8 or 9");
        if (split < 0) throw Error(GIMIC_EINVAL, "split count must be >= 0");
        c->mixed_variant = variant == 0 ? -1 : variant;
        c->mixed_split = variant >= 3 ? split : 0;
        for (auto &e : c->tune_ms) e = 0.f;
    });
}

int gimic_ctx_set_fp64_radius(gimic_ctx *c, double radius) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        if (!(radius >= 0.0) || !std::isfinite(radius))
            throw Error(GIMIC_EINVAL, "radius must be finite and >= 0");
        c->fp64_radius = radius;
    });
}

int gimic_ctx_set_chunk_bytes(gimic_ctx *c, int64_t bytes) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        if (bytes < -1) throw Error(GIMIC_EINVAL, "chunk bytes must be >= -1");
        c->chunk_bytes = bytes == 0    ? kChunkBytesDefault
                         : bytes == -1 ? c->chunk_bytes_large
                                       : (size_t)bytes;
    });
}

int gimic_ctx_chunk_bytes(const gimic_ctx *c, int64_t *bytes) {
    return guarded([&] {
        if (!c || !bytes) throw Error(GIMIC_EINVAL, "null argument");
        *bytes = (int64_t)c->chunk_bytes;
    });
}

int gimic_ctx_fp64_refined(const gimic_ctx *c, double *radius, int64_t *npts) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        if (radius) *radius = c->fp64_radius;
        if (npts) *npts = c->refined_last;
    });
}

int gimic_ctx_set_tile_stats(gimic_ctx *c, int on) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        c->tile_stats = on != 0;
        if (c->tile_stats)
            for (auto &v : c->tile_acc) v = 0.0;
    });
}

int gimic_ctx_tile_stats(const gimic_ctx *c, double *out) {
    return guarded([&] {
        if (!c || !out) throw Error(GIMIC_EINVAL, "null argument");
        for (int k = 0; k < 4; k++) out[k] = c->tile_acc[k];
        out[4] = c->lda / 32;
        out[5] = c->tile_acc[4];
        out[6] = c->tile_acc[5];
    });
}

int gimic_ctx_set_graph(gimic_ctx *c, int enable) {
    return guarded([&] {
        if (!c) throw Error(GIMIC_EINVAL, "null context");
        c->graphs = enable != 0;
    });
}

int gimic_ctx_graph_stats(const gimic_ctx *c, long long *builds, long long *launches) {
    return guarded([&] {
        if (!c || !builds || !launches) throw Error(GIMIC_EINVAL, "null argument");
        *builds = c->graph_builds;
        *launches = c->graph_launches;
    });
}

int gimic_ctx_properties(gimic_ctx *c, int64_t npts, const double *r, const double *w, int npart,
                         const int64_t *part_offs, int nnuc, const double *nuclei, int spin,
                         int precision, double *shield, double *magnet, double *jt) {
    return guarded([&] {
        if (!c || !r || !w || !part_offs || !magnet || npts < 0 || npart < 1 || nnuc < 0 ||
            (nnuc > 0 && (!nuclei || !shield)))
            throw Error(GIMIC_EINVAL, "null or invalid argument");
        if (part_offs[0] != 0 || part_offs[npart] != npts)
            throw Error(GIMIC_EINVAL, "part_offs must run from 0 to npts");
        if (nnuc >= 65535) throw Error(GIMIC_EINVAL, "at most 65534 nuclei");
        for (int k = 0; k < npart; k++)
            if (part_offs[k + 1] < part_offs[k])
                throw Error(GIMIC_EINVAL, "part_offs must be non-decreasing");
        hipStream_t st = c->stream;
        const size_t np = (size_t)std::max<int64_t>(npts, 1);
        c->prop_in.alloc(np * 32, st, false);  // points then weights
        c->prop_jt.alloc(np * 72, st, false);
        c->prop_offs.alloc((size_t)(npart + 1) * 8, st, false);
        c->prop_nuc.alloc((size_t)std::max(nnuc, 1) * 24, st, false);
        c->prop_out.alloc((size_t)(nnuc + 1) * npart * 40, st, false);
        double *d_r = c->prop_in.as<double>(), *d_w = d_r + 3 * np;
        if (npts > 0) {
            hip_check(hipMemcpyAsync(d_r, r, (size_t)npts * 24, hipMemcpyHostToDevice, st), "copy r");
            hip_check(hipMemcpyAsync(d_w, w, (size_t)npts * 8, hipMemcpyHostToDevice, st), "copy w");
        }
        hip_check(hipMemcpyAsync(c->prop_offs.p, part_offs, (size_t)(npart + 1) * 8,
                                 hipMemcpyHostToDevice, st), "copy offs");
        if (nnuc > 0)
            hip_check(hipMemcpyAsync(c->prop_nuc.p, nuclei, (size_t)nnuc * 24,
                                     hipMemcpyHostToDevice, st), "copy nuclei");
        evaluate(c, npts, d_r, c->prop_jt.as<double>(), {spin}, precision, st, nullptr);
        hip_check(launch_properties(c->prop_jt.as<double>(), d_r, d_w,
                                    (const long long *)c->prop_offs.p, c->prop_nuc.as<double>(),
                                    nnuc, npart, c->prop_out.as<double>(), st),
                  "properties");
        std::vector<double> res((size_t)(nnuc + 1) * npart * 5);
        hip_check(hipMemcpyAsync(res.data(), c->prop_out.p, res.size() * 8, hipMemcpyDeviceToHost,
                                 st), "copy out");
        if (jt && npts > 0)
            hip_check(hipMemcpyAsync(jt, c->prop_jt.p, (size_t)npts * 72, hipMemcpyDeviceToHost,
                                     st), "copy jt");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        const size_t ns = (size_t)nnuc * npart * 5;
        if (nnuc > 0) std::memcpy(shield, res.data(), ns * 8);
        std::memcpy(magnet, res.data() + ns, (size_t)npart * 5 * 8);
    });
}

int gimic_sph_transform(int l, double *T) {
    return guarded([&] {
        if (!T || l < 0 || l > kMaxL) throw Error(GIMIC_EINVAL, "bad argument");
        if (l <= 1) {  // s and p functions are used as they are (x, y, z)
            const int n = ncart(l);
            for (int i = 0; i < n * n; i++) T[i] = (i % (n + 1) == 0) ? 1.0 : 0.0;
            return;
        }
        auto v = sph_transform(l);
        std::memcpy(T, v.data(), v.size() * sizeof(double));
    });
}

int gimic_selftest_mfma32_layout(float *out) {
    return guarded([&] {
        if (!out) throw Error(GIMIC_EINVAL, "null argument");
        hip_check(run_mfma32_layout_probe(out), "mfma32 layout probe");
    });
}

int gimic_selftest_mfma_layout(double *f64, float *bf16) {
    return guarded([&] {
        if (!f64 || !bf16) throw Error(GIMIC_EINVAL, "null argument");
        hip_check(run_mfma_layout_probe(f64, bf16), "mfma layout probe");
    });
}

int gimic_quadrature(double a, double b, int order, int npts, int kind, double *pts,
                     double *wgts) {
    return guarded([&] {
        if (!pts || !wgts || npts < 1 || order < 1 || (kind != 0 && kind != 1))
            throw Error(GIMIC_EINVAL, "bad argument");
        if (npts > 1 && npts % order != 0)
            throw Error(GIMIC_EINVAL, "the point count must be a multiple of the quadrature order");
        setup_gauss_data(a, b, order, npts, kind == 1, pts, wgts);
    });
}

// ---------------- drop-in (gimic_interface.f90) ----------------

void gimic_init(const char *molfile, const char *densfile) {
    if (g_ctx) {
        delete g_ctx;
        g_ctx = nullptr;
    }
    g_spin = GIMIC_SPIN_TOTAL;
    g_magnet[0] = g_magnet[1] = g_magnet[2] = 0.0;
    if (gimic_ctx_create_files(&g_ctx, molfile, densfile, g_uhf_next, g_thrs_next) != GIMIC_OK) {
        g_ctx = nullptr;
        report("gimic_init");
    }
}

void gimic_finalize(void) {
    delete g_ctx;
    g_ctx = nullptr;
}

void gimic_set_uhf(int *uhf) { g_uhf_next = (uhf && *uhf != 0) ? 1 : 0; }

void gimic_set_magnet(const double *b) {
    if (b) std::memcpy(g_magnet, b, sizeof g_magnet);
}

void gimic_set_spin(const char *s) {
    const std::string v = s ? s : "";
    if (v == "alpha") g_spin = GIMIC_SPIN_ALPHA;
    else if (v == "beta") g_spin = GIMIC_SPIN_BETA;
    else if (v == "total") g_spin = GIMIC_SPIN_TOTAL;
    else if (v == "spindens") g_spin = GIMIC_SPIN_SPINDENS;
    else {
        g_error = "Invalid spin case.";
        report("gimic_set_spin");
    }
}

void gimic_set_screening(const double *thrs) {
    if (thrs) g_thrs_next = *thrs;
}

void gimic_calc_jtensor(const double *r, double *jt) {
    if (!jt) return;
    if (!g_ctx) {
        g_error = "gimic_init has not been called";
        report("gimic_calc_jtensor");
        for (int i = 0; i < 9; i++) jt[i] = 0.0;
        return;
    }
    if (gimic_jtensor_batch(g_ctx, 1, r, jt, g_spin, GIMIC_PREC_FP64) != GIMIC_OK) {
        report("gimic_calc_jtensor");
        for (int i = 0; i < 9; i++) jt[i] = 0.0;
    }
}

void gimic_calc_jvector(const double *r, double *jv) {
    double jt[9];
    gimic_calc_jtensor(r, jt);
    if (!jv) return;
    // jv = matmul(reshape(j,(3,3)), B) (jtensor.F90:125-135)
    for (int m = 0; m < 3; m++)
        jv[m] = jt[m] * g_magnet[0] + jt[m + 3] * g_magnet[1] + jt[m + 6] * g_magnet[2];
}

void gimic_calc_modj(const double *r, double *mj) {
    double jv[3];
    gimic_calc_jvector(r, jv);
    if (mj) *mj = std::sqrt(jv[0] * jv[0] + jv[1] * jv[1] + jv[2] * jv[2]);
}

void gimic_get_gauss_points(double *a, double *b, int *npts, int *order, double *pts,
                            double *wgts) {
    if (!a || !b || !npts || !order) return;
    if (gimic_quadrature(*a, *b, *order, *npts, 0, pts, wgts) != GIMIC_OK)
        report("gimic_get_gauss_points");
}

void mkgausspoints(double *a, double *b, int *npts, int *order, double *pts, double *wgts) {
    gimic_get_gauss_points(a, b, npts, order, pts, wgts);
}

}  // extern "C"
