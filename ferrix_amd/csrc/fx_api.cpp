// fx_api.cpp -- host runtime behind include/ferrix_hip.h.
//
// Context and mesh lifetime, the host symbolic phase (structural pattern,
// element-to-slot contribution lists), the C entry points and the step
// driver restated from src/solver/step/static_step.rs:47-267.
#include "fx_internal.h"
#include "fx_newton.h"

#include <chrono>
#include <climits>
#include <cstdlib>
#include <exception>
#include <new>
#include <numeric>

namespace fx {
thread_local int64_t t_host_syncs = 0;
}

using fx::DBuf;
using fx::Error;

namespace {

// The last failure message is kept per calling thread (like errno): one
// context may be driven from several threads (admission, batch), and a
// message must not be overwritten under a reader.
thread_local std::string g_last_error;

fx_status fail_ctx(fx_ctx *, fx_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define FX_API_BEGIN try {
#define FX_API_END(ctx)                                                                                  \
    }                                                                                                    \
    catch (const Error &e) {                                                                             \
        return fail_ctx(ctx, e.code, e.msg);                                                             \
    }                                                                                                    \
    catch (const std::bad_alloc &) {                                                                     \
        return fail_ctx(ctx, FX_ERR_OOM, "host out of memory");                                          \
    }                                                                                                    \
    catch (const std::exception &e) {                                                                    \
        return fail_ctx(ctx, FX_ERR_INVALID, e.what());                                                  \
    }                                                                                                    \
    return FX_OK;

void require(bool ok, const std::string &msg) {
    if (!ok)
        throw Error{FX_ERR_INVALID, msg};
}

using clk = std::chrono::steady_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

// ---------------------------------------------------------------------------
// symbolic phase
// ---------------------------------------------------------------------------
struct Symbolic {
    std::vector<int32_t> nb_ptr, nb_col, blk_row, blk_cptr, blk_cbase, nod_cptr, nod_cbase, diag_slot;
};

// Elements are visited in the given order.  Contribution bases point into
// the element-contribution buffers: ke_base[e] (+ (a*nn+b)*9), fe_base[e]
// (+ 3a).
// Sliced-ELL structure (fx_sell.h) for the batched solves of meshes whose
// node rows are short (mean <= 32 block columns: C3D4): one thread per node
// row streams the operator with fully coalesced loads, which pays when many
// systems fill the GPU.  A single solve (fx_solve) keeps the lane-group SpMV
// on the node-block layout, which has the shorter per-thread dependency
// chains and wins when one system has to hide its own latency; longer rows
// (C3D20) use lane groups everywhere.  FX_SELL=0 disables it.  Both sliced
// operators are optional: a mesh whose block columns or record offsets do
// not fit their int32 indices keeps the node-block operator (no sliced ELL)
// or the full FP32 one (no symmetric operator) instead of failing.
void build_sell(fx_mesh *m, hipStream_t s) {
    const char *e = std::getenv("FX_SELL");
    m->sell = m->n_nodes > 0 && (double)m->nblk <= 32.0 * (double)m->n_nodes && !(e && e[0] == '0');
    m->ssym = false;
    if (!m->sell)
        return;
    const int64_t n = m->n_nodes, ns = (n + fx::kSliceRows - 1) / fx::kSliceRows;
    const std::vector<int32_t> &bp = m->h_nb_ptr, &bc = m->h_nb_col;
    std::vector<int32_t> slp(ns + 1, 0);
    for (int64_t sl = 0; sl < ns; ++sl) {
        int32_t w = 0;
        for (int64_t a = sl * fx::kSliceRows; a < std::min(n, (sl + 1) * fx::kSliceRows); ++a)
            w = std::max(w, bp[a + 1] - bp[a]);
        if ((int64_t)slp[sl] + w > INT32_MAX) {
            m->sell = false;  // too many block columns for int32 slice pointers
            return;
        }
        slp[sl + 1] = slp[sl] + w;
    }
    std::vector<int32_t> sc((size_t)slp[ns] * fx::kSliceRows);
    for (int64_t sl = 0; sl < ns; ++sl)
        for (int32_t k = 0; k < slp[sl + 1] - slp[sl]; ++k)
            for (int lane = 0; lane < fx::kSliceRows; ++lane) {
                const int64_t a = sl * fx::kSliceRows + lane;
                int32_t c = 0;
                if (a < n)
                    c = k < bp[a + 1] - bp[a] ? bc[bp[a] + k] : (int32_t)a;
                sc[((size_t)slp[sl] + k) * fx::kSliceRows + lane] = c;
            }
    m->sl_ptr.upload(slp, s);
    m->scol.upload(sc, s);

    // symmetric part: diagonal position of every row, upper/lower widths
    std::vector<int32_t> d(n);
    for (int64_t a = 0; a < n; ++a) {
        const int32_t *c0 = bc.data() + bp[a], *c1 = bc.data() + bp[a + 1];
        const int32_t *it = std::lower_bound(c0, c1, (int32_t)a);
        if (it == c1 || *it != (int32_t)a)
            return;  // a row without its diagonal block (build_symbolic keeps one in every row)
        d[a] = (int32_t)(it - c0);
    }
    std::vector<int32_t> su(ns + 1, 0), sl(ns + 1, 0);
    for (int64_t q = 0; q < ns; ++q) {
        int32_t wu = 0, wl = 0;
        for (int64_t a = q * fx::kSliceRows; a < std::min(n, (q + 1) * fx::kSliceRows); ++a) {
            wu = std::max(wu, bp[a + 1] - bp[a] - d[a]);
            wl = std::max(wl, d[a]);
        }
        su[q + 1] = su[q] + wu;
        sl[q + 1] = sl[q] + wl;
        if ((int64_t)su[q + 1] * 9 * fx::kSliceRows + 9 * fx::kSliceRows > INT32_MAX)
            return;  // record offsets past int32: the full FP32 operator serves this mesh
    }
    const int32_t zero_rec = su[ns] * 9 * fx::kSliceRows;  // the zero record after the last block column
    std::vector<int32_t> scu((size_t)su[ns] * fx::kSliceRows), lr((size_t)sl[ns] * fx::kSliceRows),
        lc((size_t)sl[ns] * fx::kSliceRows), lb((size_t)sl[ns] * fx::kSliceRows);
    int64_t nu = 0, nl = 0;
    for (int64_t q = 0; q < ns; ++q)
        for (int lane = 0; lane < fx::kSliceRows; ++lane) {
            const int64_t a = q * fx::kSliceRows + lane;
            const int32_t deg = a < n ? bp[a + 1] - bp[a] : 0, da = a < n ? d[a] : 0;
            for (int32_t k = 0; k < su[q + 1] - su[q]; ++k) {
                int32_t c = a < n ? (int32_t)a : 0;
                if (a < n && da + k < deg) {
                    c = bc[bp[a] + da + k];
                    ++nu;
                }
                scu[((size_t)su[q] + k) * fx::kSliceRows + lane] = c;
            }
            for (int32_t j = 0; j < sl[q + 1] - sl[q]; ++j) {
                const size_t at = ((size_t)sl[q] + j) * fx::kSliceRows + lane;
                lr[at] = zero_rec;
                lc[at] = a < n ? (int32_t)a : 0;
                lb[at] = -1;
                if (a < n && j < da) {
                    const int32_t c = bc[bp[a] + j];  // c < a; the block (c, a) is in c's upper part
                    const int32_t *c0 = bc.data() + bp[c], *c1 = bc.data() + bp[c + 1];
                    const int32_t t = (int32_t)(std::lower_bound(c0, c1, (int32_t)a) - c0);
                    if (c0 + t == c1 || c0[t] != (int32_t)a || t < d[c])
                        return;  // not structurally symmetric (element adjacency always is)
                    const int64_t qc = c / fx::kSliceRows;
                    lr[at] = (su[qc] + (t - d[c])) * 9 * fx::kSliceRows + (int32_t)(c % fx::kSliceRows);
                    lc[at] = c;
                    lb[at] = bp[c] + t;
                    ++nl;
                }
            }
        }
    m->nblk_u = nu;
    m->nblk_l = nl;
    m->su_cols = su[ns];
    m->ndiag.upload(d, s);
    m->su_ptr.upload(su, s);
    m->sll_ptr.upload(sl, s);
    m->scol_u.upload(scu, s);
    m->lrec.upload(lr, s);
    m->lcol.upload(lc, s);
    m->lblk.upload(lb, s);
    m->ssym = true;
}

Symbolic build_symbolic(int64_t n_nodes, int64_t n_el, const int64_t *el_ptr, const int32_t *conn,
                        const std::vector<int64_t> &ke_base, const std::vector<int64_t> &fe_base) {
    Symbolic S;
    // node adjacency rows (every node keeps its own diagonal block)
    std::vector<int64_t> cnt(n_nodes + 1, 0);
    for (int64_t a = 0; a < n_nodes; ++a)
        cnt[a + 1] = 1;
    for (int64_t e = 0; e < n_el; ++e) {
        const int64_t nn = el_ptr[e + 1] - el_ptr[e];
        for (int64_t i = el_ptr[e]; i < el_ptr[e + 1]; ++i)
            cnt[conn[i] + 1] += nn;
    }
    for (int64_t a = 0; a < n_nodes; ++a)
        cnt[a + 1] += cnt[a];
    std::vector<int32_t> cols(cnt[n_nodes]);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t a = 0; a < n_nodes; ++a)
        cols[pos[a]++] = (int32_t)a;
    for (int64_t e = 0; e < n_el; ++e)
        for (int64_t i = el_ptr[e]; i < el_ptr[e + 1]; ++i)
            for (int64_t j = el_ptr[e]; j < el_ptr[e + 1]; ++j)
                cols[pos[conn[i]]++] = conn[j];
    S.nb_ptr.assign(n_nodes + 1, 0);
    std::vector<int32_t> out;
    out.reserve(cols.size() / 2 + 16);
    for (int64_t a = 0; a < n_nodes; ++a) {
        auto b = cols.begin() + cnt[a], e = cols.begin() + cnt[a + 1];
        std::sort(b, e);
        auto last = std::unique(b, e);
        out.insert(out.end(), b, last);
        if (out.size() > (size_t)INT32_MAX / 9)
            throw Error{FX_ERR_INVALID, "structural pattern exceeds 2^31 nonzeros"};
        S.nb_ptr[a + 1] = (int32_t)out.size();
    }
    S.nb_col = std::move(out);
    const int64_t nblk = S.nb_col.size();
    auto block_of = [&](int32_t ga, int32_t gb) -> int64_t {
        auto b = S.nb_col.begin() + S.nb_ptr[ga], e = S.nb_col.begin() + S.nb_ptr[ga + 1];
        return (int64_t)(std::lower_bound(b, e, gb) - S.nb_col.begin());
    };
    // block contribution lists in element order
    std::vector<int64_t> bc(nblk + 1, 0);
    std::vector<int32_t> blk_of_pair;
    int64_t npairs = 0;
    for (int64_t e = 0; e < n_el; ++e) {
        const int64_t nn = el_ptr[e + 1] - el_ptr[e];
        npairs += nn * nn;
    }
    if (npairs > INT32_MAX)
        throw Error{FX_ERR_INVALID, "too many element contributions for 32-bit lists"};
    blk_of_pair.resize(npairs);
    int64_t q = 0;
    for (int64_t e = 0; e < n_el; ++e)
        for (int64_t i = el_ptr[e]; i < el_ptr[e + 1]; ++i)
            for (int64_t j = el_ptr[e]; j < el_ptr[e + 1]; ++j) {
                int64_t blk = block_of(conn[i], conn[j]);
                blk_of_pair[q++] = (int32_t)blk;
                bc[blk + 1]++;
            }
    for (int64_t b = 0; b < nblk; ++b)
        bc[b + 1] += bc[b];
    S.blk_cptr.assign(bc.begin(), bc.end());
    S.blk_row.resize(nblk);
    for (int64_t a = 0; a < n_nodes; ++a)
        for (int32_t b = S.nb_ptr[a]; b < S.nb_ptr[a + 1]; ++b)
            S.blk_row[b] = (int32_t)a;
    S.blk_cbase.resize(npairs);
    std::vector<int64_t> fill(bc.begin(), bc.end() - 1);
    q = 0;
    for (int64_t e = 0; e < n_el; ++e) {
        const int64_t nn = el_ptr[e + 1] - el_ptr[e];
        for (int64_t a = 0; a < nn; ++a)
            for (int64_t b = 0; b < nn; ++b) {
                int64_t v = ke_base[e] + (a * nn + b) * 9;
                if (v > INT32_MAX)
                    throw Error{FX_ERR_INVALID, "element contribution buffer exceeds 32-bit indexing"};
                S.blk_cbase[fill[blk_of_pair[q++]]++] = (int32_t)v;
            }
    }
    // node lists for the internal force
    std::vector<int64_t> nc(n_nodes + 1, 0);
    for (int64_t e = 0; e < n_el; ++e)
        for (int64_t i = el_ptr[e]; i < el_ptr[e + 1]; ++i)
            nc[conn[i] + 1]++;
    for (int64_t a = 0; a < n_nodes; ++a)
        nc[a + 1] += nc[a];
    S.nod_cptr.assign(nc.begin(), nc.end());
    S.nod_cbase.resize(nc[n_nodes]);
    std::vector<int64_t> nf(nc.begin(), nc.end() - 1);
    for (int64_t e = 0; e < n_el; ++e)
        for (int64_t i = el_ptr[e]; i < el_ptr[e + 1]; ++i)
            S.nod_cbase[nf[conn[i]]++] = (int32_t)(fe_base[e] + 3 * (i - el_ptr[e]));
    // diagonal slots
    S.diag_slot.resize(3 * n_nodes);
    for (int64_t a = 0; a < n_nodes; ++a) {
        const int64_t b0 = S.nb_ptr[a], deg = S.nb_ptr[a + 1] - b0;
        const int64_t k = block_of((int32_t)a, (int32_t)a) - b0;
        for (int p = 0; p < 3; ++p)
            S.diag_slot[3 * a + p] = (int32_t)(9 * b0 + p * 3 * deg + 3 * k + p);
    }
    return S;
}

void set_device(fx_ctx *ctx) { FX_HIP(hipSetDevice(ctx->device)); }

// ---------------------------------------------------------------------------
// amplitudes (amplitude.rs:39-98)
// ---------------------------------------------------------------------------
double interp_series(double t, const double *times, const double *values, int64_t n, double sx, double sy) {
    if (n == 0)
        return 0.0;
    const double tt = t - sx;
    int64_t idx = std::lower_bound(times, times + n, tt) - times;  // partition_point(x < tt)
    if (idx == 0)
        return values[0] + sy;
    if (idx >= n)
        return values[n - 1] + sy;
    const double t0 = times[idx - 1], t1 = times[idx], v0 = values[idx - 1], v1 = values[idx];
    const double slope = (v1 - v0) / (t1 - t0);
    return v0 + slope * (tt - t0) + sy;
}

double amp_y(const fx_amplitudes &A, int a, const fx_timer &t, int origin, int current_step) {
    if (A.n > 0 && a > 0 && A.has_data[a]) {
        const int64_t b = A.ptr[a], n = A.ptr[a + 1] - A.ptr[a];
        return interp_series(A.total_time[a] ? t.global_time : t.local_time, A.times + b, A.values + b, n,
                             A.shift_x[a], A.shift_y[a]);
    }
    if (current_step > origin)
        return 1.0;
    return t.local_time / t.local_max;
}

}  // namespace

namespace fx {
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace fx

// ===========================================================================
// C ABI
// ===========================================================================
namespace fx {
constexpr size_t kPoolCap = (size_t)16 << 30;

void *pool_take(fx_ctx *ctx, size_t bytes, size_t *cap) {
    {
        std::lock_guard<std::mutex> lk(ctx->pool_mx);
        auto it = ctx->pool.lower_bound(bytes);
        if (it != ctx->pool.end() && it->first <= 2 * bytes) {
            void *p = it->second;
            *cap = it->first;
            ctx->pool_bytes -= it->first;
            ctx->pool.erase(it);
            return p;
        }
    }
    const size_t b = (bytes + 255) / 256 * 256;
    void *p = nullptr;
    FX_HIP(hipMalloc(&p, b));
    *cap = b;
    return p;
}

void pool_give(fx_ctx *ctx, void *p, size_t cap) {
    {
        std::lock_guard<std::mutex> lk(ctx->pool_mx);
        if (ctx->pool_bytes + cap <= kPoolCap) {
            ctx->pool.emplace(cap, p);
            ctx->pool_bytes += cap;
            return;
        }
    }
    (void)hipFree(p);
}

void pool_clear(fx_ctx *ctx) {
    std::lock_guard<std::mutex> lk(ctx->pool_mx);
    for (auto &kv : ctx->pool)
        (void)hipFree(kv.second);
    ctx->pool.clear();
    ctx->pool_bytes = 0;
}

constexpr size_t kPinCap = (size_t)4 << 30;

void *pin_take(fx_ctx *ctx, size_t bytes, size_t *cap) {
    {
        std::lock_guard<std::mutex> lk(ctx->pin_mx);
        auto it = ctx->pin_pool.lower_bound(bytes);
        if (it != ctx->pin_pool.end() && it->first <= 2 * bytes) {
            void *p = it->second;
            *cap = it->first;
            ctx->pin_bytes -= it->first;
            ctx->pin_pool.erase(it);
            return p;
        }
    }
    const size_t b = (bytes + 4095) / 4096 * 4096;
    void *p = nullptr;
    FX_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
    *cap = b;
    return p;
}

void pin_give(fx_ctx *ctx, void *p, size_t cap) {
    {
        std::lock_guard<std::mutex> lk(ctx->pin_mx);
        if (ctx->pin_bytes + cap <= kPinCap) {
            ctx->pin_pool.emplace(cap, p);
            ctx->pin_bytes += cap;
            return;
        }
    }
    (void)hipHostFree(p);
}

void pin_clear(fx_ctx *ctx) {
    std::lock_guard<std::mutex> lk(ctx->pin_mx);
    for (auto &kv : ctx->pin_pool)
        (void)hipHostFree(kv.second);
    ctx->pin_pool.clear();
    ctx->pin_bytes = 0;
}
}  // namespace fx

extern "C" {

const char *fx_version(void) { return "ferrix_hip 0.1 (gfx950)"; }

const char *fx_last_error(const fx_ctx *) { return g_last_error.c_str(); }

fx_status fx_ctx_memory(fx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes, int64_t *pool_bytes) {
    if (!ctx)
        return fail_ctx(nullptr, FX_ERR_INVALID, "context is NULL");
    FX_API_BEGIN
    set_device(ctx);
    size_t fr = 0, tot = 0;
    FX_HIP(hipMemGetInfo(&fr, &tot));
    if (free_bytes)
        *free_bytes = (int64_t)fr;
    if (total_bytes)
        *total_bytes = (int64_t)tot;
    if (pool_bytes) {
        std::lock_guard<std::mutex> lk(ctx->pool_mx);
        *pool_bytes = (int64_t)ctx->pool_bytes;
    }
    FX_API_END(ctx)
}

fx_status fx_ctx_create(int device, fx_ctx **out) {
    fx_ctx *ctx = nullptr;
    FX_API_BEGIN
    require(out != nullptr, "fx_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw Error{FX_ERR_NO_DEVICE, "no HIP device available"};
    if (device < 0 || device >= count)
        throw Error{FX_ERR_NO_DEVICE, "device index out of range"};
    hipDeviceProp_t prop;
    FX_HIP(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
        throw Error{FX_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", the kernels target gfx950"};
    ctx = new fx_ctx();
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    FX_HIP(hipSetDevice(device));
    fx::csr_scratch_setup(device);
    FX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    fx::GaussTable t4 = fx::gauss_table(4), t20 = fx::gauss_table(20);
    auto &T = ctx->tables;
    T.n4.upload(t4.N, ctx->stream);
    T.dn4.upload(t4.dN, ctx->stream);
    T.w4.upload(t4.w, ctx->stream);
    T.n20.upload(t20.N, ctx->stream);
    T.dn20.upload(t20.dN, ctx->stream);
    T.w20.upload(t20.w, ctx->stream);
    T.dn20n.upload(fx::c3d20_node_dn(), ctx->stream);
    FX_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx;
    ctx = nullptr;
    }
    catch (const Error &e) {
        if (ctx) {
            if (ctx->stream)
                (void)hipStreamDestroy(ctx->stream);
            if (ctx->copy_stream)
                (void)hipStreamDestroy(ctx->copy_stream);
            delete ctx;
        }
        return fail_ctx(nullptr, e.code, e.msg);
    }
    catch (const std::exception &e) {
        delete ctx;
        return fail_ctx(nullptr, FX_ERR_INVALID, e.what());
    }
    return FX_OK;
}

void fx_ctx_destroy(fx_ctx *ctx) {
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    std::lock_guard<std::mutex> lk(ctx->reg);
    for (fx_batch *b : ctx->batches)
        fx::batch_free(b);
    ctx->batches.clear();
    for (fx_mesh *m : ctx->meshes) {
        fx::batch_drop_mesh_foreign(m);  // slots in batches of other contexts
        delete m;
    }
    ctx->meshes.clear();
    fx::pool_clear(ctx);  // after the batches, whose slots return their blocks
    fx::pin_clear(ctx);
    for (auto *b : {&ctx->tables.n4, &ctx->tables.dn4, &ctx->tables.w4, &ctx->tables.n20, &ctx->tables.dn20,
                    &ctx->tables.w20, &ctx->tables.dn20n})
        b->release();
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->copy_stream)
        (void)hipStreamDestroy(ctx->copy_stream);
    delete ctx;
}

fx_status fx_mesh_upload(fx_ctx *ctx, int64_t n_nodes, const double *xyz, const double *node_temp, int64_t n_el,
                         const int64_t *el_ptr, const int32_t *conn, const int32_t *mat,
                         const fx_material_table *mats, fx_mesh **out) {
    if (!ctx)
        return fail_ctx(nullptr, FX_ERR_INVALID, "fx_mesh_upload: ctx is NULL");
    fx_mesh *m = nullptr;
    FX_API_BEGIN
    require(out && xyz && node_temp && el_ptr && conn && mat && mats, "fx_mesh_upload: NULL argument");
    *out = nullptr;
    require(n_nodes > 0 && n_nodes <= INT32_MAX / 3, "fx_mesh_upload: n_nodes out of range");
    require(n_el >= 0, "fx_mesh_upload: n_el < 0");
    require(el_ptr[0] == 0, "fx_mesh_upload: el_ptr[0] != 0");
    require(mats->n_mat >= 1, "fx_mesh_upload: no material");
    for (int64_t e = 0; e < n_el; ++e) {
        const int64_t nn = el_ptr[e + 1] - el_ptr[e];
        require(nn == 4 || nn == 20, "element " + std::to_string(e) + ": only C3D4 (4 nodes) and C3D20 (20 nodes)");
        for (int64_t i = el_ptr[e]; i < el_ptr[e + 1]; ++i)
            require(conn[i] >= 0 && conn[i] < n_nodes, "element " + std::to_string(e) + ": node index out of range");
        require(mat[e] >= 0 && mat[e] < mats->n_mat, "element " + std::to_string(e) + " has no material assigned.");
    }
    for (int k = 0; k < mats->n_mat; ++k)
        require(mats->n_e[k] <= FX_LUT_MAX && mats->n_nu[k] <= FX_LUT_MAX && mats->n_e[k] >= 0 && mats->n_nu[k] >= 0,
                "material LUT size out of range");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    m = new fx_mesh();
    m->ctx = ctx;
    m->n_nodes = n_nodes;
    m->n_dof = 3 * n_nodes;
    m->n_el = n_el;
    // element blocks by type (order kept within a type), contribution bases
    std::vector<int32_t> c4, c20, m4, m20;
    std::vector<int64_t> g4, g20, ke_base(n_el), fe_base(n_el);
    for (int64_t e = 0; e < n_el; ++e) {
        const int64_t nn = el_ptr[e + 1] - el_ptr[e];
        auto &c = nn == 4 ? c4 : c20;
        c.insert(c.end(), conn + el_ptr[e], conn + el_ptr[e + 1]);
        (nn == 4 ? m4 : m20).push_back(mat[e]);
        (nn == 4 ? g4 : g20).push_back(e);
    }
    m->blk4.nn = 4;
    m->blk4.n_el = (int64_t)g4.size();
    m->blk20.nn = 20;
    m->blk20.n_el = (int64_t)g20.size();
    m->blk4.ke_offset = 0;
    m->blk20.ke_offset = m->blk4.n_el * 144;
    m->blk4.fe_offset = 0;
    m->blk20.fe_offset = m->blk4.n_el * 12;
    m->ke_total = m->blk20.ke_offset + m->blk20.n_el * 3600;
    m->fe_total = m->blk20.fe_offset + m->blk20.n_el * 60;
    require(m->ke_total < INT32_MAX, "mesh too large for 32-bit contribution indexing");
    for (size_t k = 0; k < g4.size(); ++k) {
        ke_base[g4[k]] = m->blk4.ke_offset + (int64_t)k * 144;
        fe_base[g4[k]] = m->blk4.fe_offset + (int64_t)k * 12;
    }
    for (size_t k = 0; k < g20.size(); ++k) {
        ke_base[g20[k]] = m->blk20.ke_offset + (int64_t)k * 3600;
        fe_base[g20[k]] = m->blk20.fe_offset + (int64_t)k * 60;
    }
    Symbolic S = build_symbolic(n_nodes, n_el, el_ptr, conn, ke_base, fe_base);
    m->nblk = (int64_t)S.nb_col.size();
    m->nnz_struct = 9 * m->nblk;
    m->h_nb_ptr = S.nb_ptr;
    m->h_nb_col = S.nb_col;
    // uploads
    m->xyz.upload(xyz, 3 * n_nodes, s);
    m->temp.upload(node_temp, n_nodes, s);
    std::vector<fx::MatLut> lut(mats->n_mat);
    for (int k = 0; k < mats->n_mat; ++k) {
        fx::MatLut &L = lut[k];
        memset(&L, 0, sizeof L);
        L.n_e = mats->n_e[k];
        L.n_nu = mats->n_nu[k];
        for (int j = 0; j < FX_LUT_MAX; ++j) {
            L.t_e[j] = mats->t_e[k * FX_LUT_MAX + j];
            L.v_e[j] = mats->v_e[k * FX_LUT_MAX + j];
            L.t_nu[j] = mats->t_nu[k * FX_LUT_MAX + j];
            L.v_nu[j] = mats->v_nu[k * FX_LUT_MAX + j];
        }
    }
    m->mats.upload(lut, s);
    m->n_mat = mats->n_mat;
    m->blk4.conn.upload(c4, s);
    m->blk4.mat.upload(m4, s);
    m->blk4.gpos.upload(g4, s);
    m->blk20.conn.upload(c20, s);
    m->blk20.mat.upload(m20, s);
    m->blk20.gpos.upload(g20, s);
    m->nb_ptr.upload(S.nb_ptr, s);
    m->nb_col.upload(S.nb_col, s);
    build_sell(m, s);
    m->blk_row.upload(S.blk_row, s);
    m->blk_cptr.upload(S.blk_cptr, s);
    m->blk_cbase.upload(S.blk_cbase, s);
    m->nod_cptr.upload(S.nod_cptr, s);
    m->nod_cbase.upload(S.nod_cbase, s);
    m->diag_slot.upload(S.diag_slot, s);
    // C3D4 elements per workgroup: 128 until fx_mesh_autotune_elements
    // times the others; FX_C3D4_BLOCK=64/128/256/512 sets it at upload
    if (const char *e = std::getenv("FX_C3D4_BLOCK")) {
        const int bs = std::atoi(e);
        require(bs == 64 || bs == 128 || bs == 256 || bs == 512, "FX_C3D4_BLOCK must be 64, 128, 256 or 512");
        m->c3d4_block = bs;
    }
    m->ke.alloc(std::max<int64_t>(1, m->ke_total));
    m->fe.alloc(std::max<int64_t>(1, m->fe_total));
    m->kval.alloc(m->nnz_struct);
    m->keep.alloc(m->nnz_struct);
    FX_HIP(hipMemsetAsync(m->kval.p, 0, m->kval.bytes(), s));
    FX_HIP(hipMemsetAsync(m->keep.p, 0, m->keep.bytes(), s));
    m->inv_diag.alloc(m->n_dof);
    m->max_diag_bits.alloc(1);
    m->status.alloc(1);
    m->pcg.alloc(1);
    m->part.alloc(3 * (size_t)(m->n_dof + 8192));
    m->counters.alloc(8);
    FX_HIP(hipMemsetAsync(m->counters.p, 0, m->counters.bytes(), s));
    for (auto *b : {&m->u, &m->f_int, &m->f_ext, &m->rhs, &m->du})
        b->alloc(m->n_dof);
    FX_HIP(hipMemsetAsync(m->u.p, 0, m->u.bytes(), s));
    m->red2.alloc(2);
    fx::reset_status(m);
    FX_HIP(hipStreamSynchronize(s));
    m->device_bytes = 0;
    for (size_t b : {m->xyz.bytes(), m->temp.bytes(), m->mats.bytes(), m->blk4.conn.bytes(), m->blk20.conn.bytes(),
                     m->nb_ptr.bytes(), m->nb_col.bytes(), m->blk_row.bytes(), m->blk_cptr.bytes(), m->blk_cbase.bytes(),
                     m->nod_cptr.bytes(), m->nod_cbase.bytes(), m->diag_slot.bytes(), m->ke.bytes(), m->fe.bytes(),
                     m->kval.bytes(), m->keep.bytes(), m->part.bytes()})
        m->device_bytes += (int64_t)b;
    m->device_bytes += 13 * m->n_dof * (int64_t)sizeof(double);
    {
        std::lock_guard<std::mutex> lk(ctx->reg);
        ctx->meshes.push_back(m);
    }
    *out = m;
    m = nullptr;
    }
    catch (const Error &e) {
        delete m;
        return fail_ctx(ctx, e.code, e.msg);
    }
    catch (const std::bad_alloc &) {
        delete m;
        return fail_ctx(ctx, FX_ERR_OOM, "host out of memory");
    }
    catch (const std::exception &e) {
        delete m;
        return fail_ctx(ctx, FX_ERR_INVALID, e.what());
    }
    return FX_OK;
}

void fx_mesh_destroy(fx_mesh *m) {
    if (!m)
        return;
    fx_ctx *ctx = m->ctx;
    std::lock_guard<std::mutex> lk(ctx->reg);
    auto it = std::find(ctx->meshes.begin(), ctx->meshes.end(), m);
    if (it == ctx->meshes.end())
        return;  // not a live mesh of its context
    ctx->meshes.erase(it);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    fx::drop_poly_batch(m);
    for (fx_batch *b : ctx->batches)
        fx::batch_drop_mesh(b, m);
    fx::batch_drop_mesh_foreign(m);
    delete m;
}

fx_status fx_mesh_info(const fx_mesh *m, int64_t *n_dof, int64_t *nnz_struct, int64_t *device_bytes) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "fx_mesh_info: mesh is NULL");
    if (n_dof)
        *n_dof = m->n_dof;
    if (nnz_struct)
        *nnz_struct = m->nnz_struct;
    if (device_bytes)
        *device_bytes = m->device_bytes;
    return FX_OK;
}

fx_status fx_state_set_u(fx_mesh *m, const double *u) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    set_device(m->ctx);
    if (u)
        FX_HIP(hipMemcpyAsync(m->u.p, u, m->n_dof * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
    else
        fx::launch_fill(m, m->u.p, m->n_dof, 0.0);
    FX_HIP(hipStreamSynchronize(m->ctx->stream));
    FX_API_END(m->ctx)
}

fx_status fx_state_get_u(fx_mesh *m, double *u) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(u != nullptr, "fx_state_get_u: u is NULL");
    set_device(m->ctx);
    FX_HIP(hipMemcpyAsync(u, m->u.p, m->n_dof * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
    FX_HIP(hipStreamSynchronize(m->ctx->stream));
    FX_API_END(m->ctx)
}

fx_status fx_profile_spmv(fx_mesh *m, int32_t precision, int32_t reps, int32_t spmv_group, double *ms_per_launch,
                          double *bytes_per_launch) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(m->k_valid, "fx_profile_spmv: assemble first");
    require(reps > 0 && ms_per_launch && bytes_per_launch, "fx_profile_spmv: bad argument");
    set_device(m->ctx);
    *ms_per_launch = fx::profile_spmv(m, precision, reps, spmv_group, bytes_per_launch);
    FX_API_END(m->ctx)
}

fx_status fx_measure_fp64_peak(fx_ctx *ctx, int32_t reps, double *valu_tflops, double *mfma_tflops) {
    if (!ctx)
        return fail_ctx(nullptr, FX_ERR_INVALID, "ctx is NULL");
    FX_API_BEGIN
    require(reps > 0 && valu_tflops && mfma_tflops, "fx_measure_fp64_peak: bad argument");
    set_device(ctx);
    fx::measure_fp64_peak(ctx, reps, valu_tflops, mfma_tflops);
    FX_API_END(ctx)
}

fx_status fx_profile_element(fx_mesh *m, int32_t reps, int32_t mfma, double *ms_per_launch,
                             double *flop_per_launch) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(reps > 0 && ms_per_launch && flop_per_launch, "fx_profile_element: bad argument");
    set_device(m->ctx);
    *ms_per_launch = fx::profile_element(m, reps, mfma != 0, flop_per_launch);
    FX_API_END(m->ctx)
}

fx_status fx_mesh_autotune_elements(fx_mesh *m, int32_t reps, int32_t allow_mfma, int32_t *c3d4_block_out,
                                    int32_t *c3d20_mfma_out, double *ms_per_launch) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(reps > 0, "fx_mesh_autotune_elements: reps must be positive");
    set_device(m->ctx);
    const double ms = fx::autotune_elements(m, reps, allow_mfma != 0);
    if (c3d4_block_out)
        *c3d4_block_out = m->c3d4_block;
    if (c3d20_mfma_out)
        *c3d20_mfma_out = m->blk20.n_el ? m->c20_mfma : -1;
    if (ms_per_launch)
        *ms_per_launch = ms;
    FX_API_END(m->ctx)
}

fx_status fx_mesh_autotune_spmv(fx_mesh *m, int32_t precision, int32_t reps, int32_t *spmv_group_out,
                                double *ms_per_launch) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(m->k_valid, "fx_mesh_autotune_spmv: assemble first");
    require(reps > 0, "fx_mesh_autotune_spmv: reps must be positive");
    require(precision == FX_PREC_FP64 || precision == FX_PREC_MIXED, "fx_mesh_autotune_spmv: unknown precision");
    set_device(m->ctx);
    const int G = fx::autotune_spmv(m, precision, reps, ms_per_launch);
    if (spmv_group_out)
        *spmv_group_out = G;
    FX_API_END(m->ctx)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// shared pieces of the assembly seam
// ---------------------------------------------------------------------------
namespace {

double assemble_k(fx_mesh *m, bool nlgeom) {
    fx::reset_status(m);
    fx::launch_element(m, true, false, nlgeom);
    fx::launch_assemble(m);
    fx::check_status(m, "Assembler::assemble");
    m->last_max_diag = fx::device_max_diag(m);
    m->k_valid = true;
    return m->last_max_diag;
}

void internal_force_dev(fx_mesh *m, bool nlgeom) {
    fx::reset_status(m);
    fx::launch_element(m, false, true, nlgeom);
    fx::launch_gather_f(m);
}

bool env_on(const char *name) {
    const char *e = std::getenv(name);
    return !(e && e[0] == '0');
}

// three timing events of a Newton iteration (front start, front end, queued solve end)
struct Events3 {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    Events3() {
        for (auto &x : e)
            FX_HIP(hipEventCreate(&x));
    }
    ~Events3() {
        for (auto &x : e)
            if (x)
                (void)hipEventDestroy(x);
    }
};

// the element failure message of check_status
std::string element_error(const char *what, int code, int64_t bad_element) {
    std::string msg = std::string(what) + ": ";
    if (code == FX_ERR_SINGULAR_JACOBIAN)
        msg += "Singular Jacobian";
    else
        msg += "element has no Young's modulus or Poisson's ratio at its temperature";
    return msg + " (element position " + std::to_string(bad_element) + ")";
}

// f_ext at the timer's time: amplitude-scaled loads summed per DOF in deck
// order (static_step.rs:111-126), scattered into a zeroed m->f_ext
void set_fext(fx_mesh *m, const fx_step_def *sd, const fx_timer &timer, int step_id,
              const std::vector<int32_t> &touched, std::vector<double> &fext_h, std::vector<double> &tval,
              const fx::DBuf<int32_t> &d_touched, fx::DBuf<double> &d_tval) {
    hipStream_t s = m->ctx->stream;
    for (int32_t d : touched)
        fext_h[d] = 0.0;
    for (int64_t k = 0; k < sd->n_loads; ++k)
        fext_h[sd->load_dof[k]] += amp_y(sd->amps, sd->load_amp[k], timer, sd->load_origin[k], step_id) * sd->load_val[k];
    for (size_t k = 0; k < touched.size(); ++k)
        tval[k] = fext_h[touched[k]];
    if (!touched.empty())
        FX_HIP(hipMemcpyAsync(d_tval.p, tval.data(), tval.size() * sizeof(double), hipMemcpyHostToDevice, s));
    fx::launch_fill(m, m->f_ext.p, m->n_dof, 0.0);
    fx::launch_scatter(m, (int64_t)touched.size(), d_touched.p, d_tval.p, m->f_ext.p);
}

// The step's linear solve through the caller's solver (fx_step_def.solve):
// rhs (m->rhs) to the host, the call, x back into m->du.
fx_status hook_solve(fx_mesh *m, const fx_step_def *sd, fx_solve_report *sr) {
    hipStream_t s = m->ctx->stream;
    const int64_t n = m->n_dof;
    std::vector<double> b(n), x(n, 0.0);
    FX_HIP(hipMemcpyAsync(b.data(), m->rhs.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    ++fx::t_host_syncs;
    int64_t its = 0;
    const int32_t rc = sd->solve(sd->solve_user, m, n, b.data(), x.data(), &its);
    *sr = fx_solve_report{};
    sr->iterations = its;
    if (rc != FX_OK)
        return (fx_status)rc;
    FX_HIP(hipMemcpyAsync(m->du.p, x.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    FX_HIP(hipStreamSynchronize(s));  // x is a local buffer
    ++fx::t_host_syncs;
    return FX_OK;
}

// BC records grouped per DOF in first-appearance order, records in deck order
struct BcGroups {
    std::vector<int32_t> ud_dof, ud_ptr, rec_order;
};

BcGroups group_bcs(int64_t n, const int64_t *dofs, int64_t n_dof) {
    BcGroups g;
    std::vector<int64_t> first(n_dof, -1);
    std::vector<std::vector<int32_t>> recs;
    for (int64_t k = 0; k < n; ++k) {
        int64_t d = dofs[k];
        if (d < 0 || d >= n_dof)
            throw Error{FX_ERR_INVALID, "boundary condition DOF out of range"};
        if (first[d] < 0) {
            first[d] = (int64_t)recs.size();
            recs.emplace_back();
            g.ud_dof.push_back((int32_t)d);
        }
        recs[first[d]].push_back((int32_t)k);
    }
    g.ud_ptr.push_back(0);
    for (auto &r : recs) {
        g.rec_order.insert(g.rec_order.end(), r.begin(), r.end());
        g.ud_ptr.push_back((int32_t)g.rec_order.size());
    }
    return g;
}

struct DevBc {
    DBuf<int32_t> ud_dof, ud_ptr, rec_order;
    DBuf<double> target;
    int64_t n_unique = 0;
    void upload(const BcGroups &g, hipStream_t s) {
        ud_dof.upload(g.ud_dof, s);
        ud_ptr.upload(g.ud_ptr, s);
        rec_order.upload(g.rec_order, s);
        n_unique = (int64_t)g.ud_dof.size();
    }
};

void check_solve_opts(const fx_solve_opts &o) {
    require(o.precision == FX_PREC_FP64 || o.precision == FX_PREC_MIXED || o.precision == FX_PREC_MIXED_POLY,
            "solve: unknown precision");
    require(o.precision != FX_PREC_MIXED_POLY || (o.poly_degree >= 0 && o.poly_degree <= fx::kPolyMax),
            "solve: poly_degree must be 0 (default 15) or 1.." + std::to_string(fx::kPolyMax));
    require(o.stop_mode == FX_STOP_ABS_SQRT_RZ || o.stop_mode == FX_STOP_REL_R2, "solve: unknown stop mode");
    require(o.max_iter >= 0, "solve: max_iter < 0");
    require(o.spmv_group == 0 || o.spmv_group == 2 || o.spmv_group == 4 || o.spmv_group == 8 ||
                o.spmv_group == 16 || o.spmv_group == 32 || o.spmv_group == 64,
            "solve: spmv_group must be 0 or a power of two in 2..64");
}

std::string solve_error(fx_status st, const fx_solve_opts &o, const fx_solve_report &r) {
    if (st == FX_ERR_NOT_CONVERGED)
        return "PCG did not converge after " + std::to_string(o.max_iter) + " iterations (criterion " +
               std::to_string(r.final_criterion) + ")";
    return "PCG breakdown: p.Ap <= 0 or not finite after " + std::to_string(r.iterations) + " iterations";
}

}  // namespace

extern "C" {

void fx_solve_opts_default(fx_solve_opts *o) {
    if (!o)
        return;
    memset(o, 0, sizeof *o);
    o->precision = FX_PREC_FP64;
    o->stop_mode = FX_STOP_ABS_SQRT_RZ;
    o->tol = 1e-8;
    o->max_iter = 1000;
    o->inner_tol = 1e-3;
    o->max_refine = 50;
    o->spmv_group = 0;
    o->poly_degree = 0;
}

fx_status fx_assemble(fx_mesh *m, int32_t nlgeom, double *max_diag_out, int64_t *nnz_out) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    set_device(m->ctx);
    double md = assemble_k(m, nlgeom != 0);
    if (max_diag_out)
        *max_diag_out = md;
    if (nnz_out)
        *nnz_out = fx::count_kept(m);
    FX_API_END(m->ctx)
}

fx_status fx_internal_force(fx_mesh *m, int32_t nlgeom, double *f_int_out) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    set_device(m->ctx);
    internal_force_dev(m, nlgeom != 0);
    fx::check_status(m, "Assembler::assemble_internal_force");
    if (f_int_out) {
        FX_HIP(hipMemcpyAsync(f_int_out, m->f_int.p, m->n_dof * sizeof(double), hipMemcpyDeviceToHost,
                              m->ctx->stream));
        FX_HIP(hipStreamSynchronize(m->ctx->stream));
    }
    FX_API_END(m->ctx)
}

fx_status fx_stress_strain(fx_mesh *m, double *stress_out, double *strain_out) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(stress_out && strain_out, "fx_stress_strain: NULL argument");
    set_device(m->ctx);
    fx::reset_status(m);
    fx::launch_recover(m);
    fx::check_status(m, "StaticStep::calculate_stress_strain");
    DBuf<double> S, E;
    S.alloc(std::max<int64_t>(1, 6 * m->n_nodes));
    E.alloc(std::max<int64_t>(1, 6 * m->n_nodes));
    fx::launch_gather_recover(m, S.p, E.p);
    hipStream_t s = m->ctx->stream;
    FX_HIP(hipMemcpyAsync(stress_out, S.p, 6 * m->n_nodes * sizeof(double), hipMemcpyDeviceToHost, s));
    FX_HIP(hipMemcpyAsync(strain_out, E.p, 6 * m->n_nodes * sizeof(double), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    FX_API_END(m->ctx)
}

fx_status fx_matrix_get(fx_mesh *m, int64_t *row_ptr, int32_t *col_idx, double *values) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(row_ptr && col_idx, "fx_matrix_get: NULL argument");
    require(m->k_valid, "fx_matrix_get: no assembled matrix");
    set_device(m->ctx);
    std::vector<uint8_t> keep(m->nnz_struct);
    std::vector<double> val(values ? m->nnz_struct : 0);
    FX_HIP(hipMemcpyAsync(keep.data(), m->keep.p, keep.size(), hipMemcpyDeviceToHost, m->ctx->stream));
    if (values)
        FX_HIP(hipMemcpyAsync(val.data(), m->kval.p, val.size() * sizeof(double), hipMemcpyDeviceToHost,
                              m->ctx->stream));
    FX_HIP(hipStreamSynchronize(m->ctx->stream));
    int64_t out = 0;
    row_ptr[0] = 0;
    for (int64_t a = 0; a < m->n_nodes; ++a) {
        const int64_t b0 = m->h_nb_ptr[a], deg = m->h_nb_ptr[a + 1] - b0;
        for (int p = 0; p < 3; ++p) {
            const int64_t base = 9 * b0 + p * 3 * deg;
            for (int64_t k = 0; k < 3 * deg; ++k) {
                if (!keep[base + k])
                    continue;
                col_idx[out] = 3 * m->h_nb_col[b0 + k / 3] + (int32_t)(k % 3);
                if (values)
                    values[out] = val[base + k];
                ++out;
            }
            row_ptr[3 * a + p + 1] = out;
        }
    }
    FX_API_END(m->ctx)
}

fx_status fx_apply_penalty(fx_mesh *m, int64_t n_bc, const int64_t *dofs, const double *targets,
                           double *rhs_inout, double *penalty_out) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(m->k_valid, "fx_apply_penalty: assemble first");
    require(n_bc == 0 || (dofs && targets), "fx_apply_penalty: NULL argument");
    set_device(m->ctx);
    hipStream_t s = m->ctx->stream;
    double pen = m->last_max_diag > 0.0 ? m->last_max_diag * fx::kPenaltyScale : 0.0;
    if (penalty_out)
        *penalty_out = pen;
    if (pen > 0.0 && n_bc > 0) {
        BcGroups g = group_bcs(n_bc, dofs, m->n_dof);
        DevBc d;
        d.upload(g, s);
        d.target.upload(targets, n_bc, s);
        if (rhs_inout)
            FX_HIP(hipMemcpyAsync(m->rhs.p, rhs_inout, m->n_dof * sizeof(double), hipMemcpyHostToDevice, s));
        fx::launch_penalty(m, d.n_unique, d.ud_dof.p, d.ud_ptr.p, d.target.p, d.rec_order.p,
                           rhs_inout ? m->rhs.p : nullptr, m->u.p);
        if (rhs_inout)
            FX_HIP(hipMemcpyAsync(rhs_inout, m->rhs.p, m->n_dof * sizeof(double), hipMemcpyDeviceToHost, s));
        FX_HIP(hipStreamSynchronize(s));
    }
    FX_API_END(m->ctx)
}

fx_status fx_solve(fx_mesh *m, const double *rhs, double *x_out, const fx_solve_opts *opts,
                   fx_solve_report *report) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(rhs && x_out, "fx_solve: NULL argument");
    require(m->k_valid, "fx_solve: assemble first");
    fx_solve_opts o;
    if (opts)
        o = *opts;
    else
        fx_solve_opts_default(&o);
    check_solve_opts(o);
    set_device(m->ctx);
    hipStream_t s = m->ctx->stream;
    FX_HIP(hipMemcpyAsync(m->rhs.p, rhs, m->n_dof * sizeof(double), hipMemcpyHostToDevice, s));
    fx_solve_report r{};
    fx_status st = fx::pcg_solve(m, m->rhs.p, m->du.p, o, &r);
    FX_HIP(hipMemcpyAsync(x_out, m->du.p, m->n_dof * sizeof(double), hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    if (report)
        *report = r;
    if (st != FX_OK)
        throw Error{st, solve_error(st, o, r)};
    FX_API_END(m->ctx)
}

fx_status fx_solve_csr(fx_ctx *ctx, int64_t n, const int64_t *row_ptr, const int64_t *col_idx,
                       const double *values, const double *rhs, double *x_out, const fx_solve_opts *opts,
                       fx_solve_report *report) {
    if (!ctx)
        return fail_ctx(nullptr, FX_ERR_INVALID, "context is NULL");
    FX_API_BEGIN
    require(n >= 0, "fx_solve_csr: n < 0");
    require(row_ptr && (n == 0 || (rhs && x_out)), "fx_solve_csr: NULL argument");
    require(row_ptr[n] == 0 || (col_idx && values), "fx_solve_csr: NULL argument");
    fx_solve_opts o;
    if (opts)
        o = *opts;
    else
        fx_solve_opts_default(&o);
    check_solve_opts(o);
    require(o.precision == FX_PREC_FP64, "fx_solve_csr: only FX_PREC_FP64 (the mixed solve needs a mesh's "
                                         "node-block operator: fx_solve)");
    if (report)
        *report = fx_solve_report{};
    if (n == 0)
        return FX_OK;
    set_device(ctx);
    fx_solve_report r{};
    const fx_status st = fx::solve_csr_host(ctx, n, row_ptr, col_idx, values, rhs, x_out, o, &r);
    if (report)
        *report = r;
    if (st != FX_OK)
        throw Error{st, solve_error(st, o, r)};
    FX_API_END(ctx)
}

fx_status fx_solve_csr_noalloc(fx_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                               const int64_t *col_dev, const double *val_dev, const double *rhs_dev, double *x_dev,
                               const fx_solve_opts *opts, fx_solve_report *report) {
    if (!ctx)
        return fail_ctx(nullptr, FX_ERR_INVALID, "context is NULL");
    FX_API_BEGIN
    require(n >= 0 && nnz >= 0, "fx_solve_csr_noalloc: n < 0 or nnz < 0");
    require(row_ptr_dev && (n == 0 || (rhs_dev && x_dev)) && (nnz == 0 || (col_dev && val_dev)),
            "fx_solve_csr_noalloc: NULL argument");
    fx_solve_opts o;
    if (opts)
        o = *opts;
    else
        fx_solve_opts_default(&o);
    check_solve_opts(o);
    require(o.precision == FX_PREC_FP64, "fx_solve_csr_noalloc: only FX_PREC_FP64");
    if (report)
        *report = fx_solve_report{};
    if (n == 0)
        return FX_OK;
    set_device(ctx);
    fx_solve_report r{};
    const fx_status st = fx::solve_csr_noalloc(ctx, n, nnz, row_ptr_dev, col_dev, val_dev, rhs_dev, x_dev, o, &r);
    if (report)
        *report = r;
    if (st != FX_OK)
        throw Error{st, solve_error(st, o, r)};
    FX_API_END(ctx)
}

// ---------------------------------------------------------------------------
// StaticStep::solve (static_step.rs:47-267)
// ---------------------------------------------------------------------------
fx_status fx_batch_add_system(fx_batch *B, fx_mesh *m, int32_t n_rhs, int64_t n_loads, const int32_t *load_case,
                              const int64_t *load_dof, const double *load_val, int64_t n_bc, const int64_t *bc_dofs,
                              const double *bc_targets, int32_t *slot_out) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(B && slot_out, "fx_batch_add_system: NULL argument");
    require(n_rhs >= 1 && n_rhs <= 4, "fx_batch_add_system: 1 to 4 right-hand sides per system");
    require(n_loads >= 0 && (n_loads == 0 || (load_case && load_dof && load_val)),
            "fx_batch_add_system: load arrays missing");
    require(n_bc >= 0 && (n_bc == 0 || (bc_dofs && bc_targets)), "fx_batch_add_system: BC arrays missing");
    set_device(m->ctx);
    hipStream_t s = m->ctx->stream;
    const int64_t n = m->n_dof;
    require(n * n_rhs < ((int64_t)1 << 31), "fx_batch_add_system: n_dof x n_rhs must stay below 2^31");
    // f_ext of each load case, static_step.rs:102-108: per (case, DOF) the
    // records' values added in record order from 0.0
    std::vector<int64_t> key((size_t)n_loads), order((size_t)n_loads);
    for (int64_t k = 0; k < n_loads; ++k) {
        require(load_case[k] >= 0 && load_case[k] < n_rhs && load_dof[k] >= 0 && load_dof[k] < n,
                "fx_batch_add_system: load record out of range");
        key[k] = (int64_t)load_case[k] * n + load_dof[k];
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    std::vector<int32_t> lkey;
    std::vector<double> lval;
    for (size_t i = 0; i < order.size();) {
        const int64_t kk = key[order[i]];
        double acc = 0.0;
        for (; i < order.size() && key[order[i]] == kk; ++i)
            acc += load_val[order[i]];
        lkey.push_back((int32_t)kk);
        lval.push_back(acc);
    }
    // the penalty records, uploaded when they differ from the last admission's
    if (m->adm_key_dof.size() != (size_t)n_bc || !std::equal(bc_dofs, bc_dofs + n_bc, m->adm_key_dof.begin()) ||
        !std::equal(bc_targets, bc_targets + n_bc, m->adm_key_val.begin())) {
        BcGroups g = group_bcs(n_bc, bc_dofs, n);
        m->adm_ud_dof.upload(g.ud_dof, s);
        m->adm_ud_ptr.upload(g.ud_ptr, s);
        m->adm_rec_order.upload(g.rec_order, s);
        m->adm_target.upload(bc_targets, (size_t)n_bc, s);
        m->adm_n_unique = (int64_t)g.ud_dof.size();
        m->adm_key_dof.assign(bc_dofs, bc_dofs + n_bc);
        m->adm_key_val.assign(bc_targets, bc_targets + n_bc);
    }
    if (m->adm_rhs_ev)  // another context's batch read the previous admission's right-hand sides
        FX_HIP(hipStreamWaitEvent(s, m->adm_rhs_ev, 0));
    m->adm_rhs.alloc((size_t)n * n_rhs);
    fx::launch_fill(m, m->adm_rhs.p, n * n_rhs, 0.0, s);
    if (!lkey.empty()) {
        m->adm_lkey.upload(lkey, s);
        m->adm_lval.upload(lval, s);
        fx::launch_scatter(m, (int64_t)lkey.size(), m->adm_lkey.p, m->adm_lval.p, m->adm_rhs.p);
    }
    // Assembler::assemble with max_diag left on the device, then the penalty
    // on K and the first right-hand side, and on the others alone
    fx::launch_reset_status(m);
    fx::launch_element(m, true, false, false);
    fx::launch_assemble(m);
    fx::launch_penalty_dev(m, m->adm_n_unique, m->adm_ud_dof.p, m->adm_ud_ptr.p, m->adm_target.p,
                           m->adm_rec_order.p, m->adm_rhs.p, m->u.p);
    for (int r = 1; r < n_rhs; ++r)
        fx::launch_penalty_dev(m, m->adm_n_unique, m->adm_ud_dof.p, m->adm_ud_ptr.p, m->adm_target.p,
                               m->adm_rec_order.p, m->adm_rhs.p + (size_t)r * n, m->u.p, true);
    m->k_valid = true;
    m->last_max_diag = std::nan("");  // on the device; read when a host path needs it
    const fx_status st = fx::batch_add_impl(B, m, n_rhs, m->adm_rhs.p, true, m->status.p, slot_out);
    if (st != FX_OK)
        return st;  // batch_add_impl set the message
    FX_API_END(m->ctx)
}

fx_status fx_static_step(fx_mesh *m, const fx_step_def *sd, fx_timer *timer, fx_step_report *report) {
    if (!m)
        return fail_ctx(nullptr, FX_ERR_INVALID, "mesh is NULL");
    FX_API_BEGIN
    require(sd && timer, "fx_static_step: NULL argument");
    require(sd->n_loads == 0 || (sd->load_dof && sd->load_val && sd->load_amp && sd->load_origin),
            "fx_static_step: load arrays missing");
    require(sd->n_bcs == 0 || (sd->bc_dof && sd->bc_val && sd->bc_amp && sd->bc_origin),
            "fx_static_step: BC arrays missing");
    const int n_amp = sd->amps.n;
    auto amp_ok = [&](int a) { return a == 0 || (a > 0 && a < n_amp); };
    for (int64_t k = 0; k < sd->n_loads; ++k)
        require(sd->load_dof[k] >= 0 && sd->load_dof[k] < m->n_dof && amp_ok(sd->load_amp[k]),
                "fx_static_step: load record out of range");
    for (int64_t k = 0; k < sd->n_bcs; ++k)
        require(sd->bc_dof[k] >= 0 && sd->bc_dof[k] < m->n_dof && amp_ok(sd->bc_amp[k]),
                "fx_static_step: boundary record out of range");
    const fx_solve_opts &opts = sd->solver == FX_SOLVER_ITERATIVE ? sd->solve_opts : sd->direct_opts;
    check_solve_opts(opts);
    set_device(m->ctx);
    hipStream_t s = m->ctx->stream;
    const int64_t n = m->n_dof;
    const bool nlgeom = sd->nlgeom != 0;
    const int step_id = sd->step_id;
    auto t_start = clk::now();
    fx_step_report rep{};

    // Step::solve: timer.new_step(time_period)
    timer->local_time = 0.0;
    timer->local_max = sd->time_period;

    // per-step constant data: BC groups, BC mask, load/target buffers
    BcGroups g = group_bcs(sd->n_bcs, sd->bc_dof, n);
    DevBc dbc;
    dbc.upload(g, s);
    std::vector<uint8_t> mask(n, 0);
    for (int64_t k = 0; k < sd->n_bcs; ++k)
        mask[sd->bc_dof[k]] = 1;
    DBuf<uint8_t> dmask;
    dmask.upload(mask, s);
    std::vector<double> targets(sd->n_bcs);
    dbc.target.alloc(std::max<int64_t>(1, sd->n_bcs));
    std::vector<double> fext_h(n, 0.0);
    std::vector<int32_t> touched;
    std::vector<char> is_touched(n, 0);
    for (int64_t k = 0; k < sd->n_loads; ++k)
        if (!is_touched[sd->load_dof[k]]) {
            is_touched[sd->load_dof[k]] = 1;
            touched.push_back((int32_t)sd->load_dof[k]);
        }
    DBuf<int32_t> d_touched;
    d_touched.upload(touched, s);
    DBuf<double> d_tval, u_start, norms;
    d_tval.alloc(std::max<size_t>(1, touched.size()));
    u_start.alloc(n);
    norms.alloc(2);
    std::vector<double> tval(touched.size());
    FX_HIP(hipMemcpyAsync(u_start.p, m->u.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    // The Newton iteration on the device (fx_newton.hip) unless FX_NR_DEVICE=0;
    // its front as a replayed graph unless FX_NR_GRAPH=0 (plain launches)
    const bool dev_nr = env_on("FX_NR_DEVICE"), use_graph = env_on("FX_NR_GRAPH");
    // the solve outside the device front: the mixed solves (host-driven
    // refinement, or the polynomial batch) and a caller's solver
    const bool host_solve = opts.precision != FX_PREC_FP64 || sd->solve != nullptr;
    fx::NewtonGraph ng(m);
    Events3 ev;
    bool nr_graph_ready = false;
    // SpMVs of the FP64 solves by NR index: the previous increment's and this one's
    std::vector<int64_t> iters_prev, iters_cur;
    const int G = fx::pcg_group(m, opts);
    if (dev_nr && !host_solve)
        fx::pcg_prepare(m);
    const int64_t syncs0 = fx::t_host_syncs;

    double current_time = 0.0;
    int64_t n_inc = 0;
    double dt = sd->initial_time_increment;
    while (current_time < sd->time_period) {
        ++n_inc;
        if (n_inc > sd->max_increments)
            throw Error{FX_ERR_NOT_CONVERGED,
                        "Maximum increment count of " + std::to_string(sd->max_increments) + " exceeded."};
        if (current_time + dt > sd->time_period)
            dt = sd->time_period - current_time;
        timer->global_time += dt;  // timer.new_increment, never rolled back
        timer->local_time += dt;
        rep.increments++;
        FX_HIP(hipMemcpyAsync(m->u.p, u_start.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        bool converged = false;
        if (dev_nr) {
            // The device-resident Newton iteration (fx_newton.hip).  The
            // loads and BC targets depend only on the increment's time:
            // formed once here (the host loop forms the same values every
            // iteration).
            set_fext(m, sd, *timer, step_id, touched, fext_h, tval, d_touched, d_tval);
            if (sd->n_bcs > 0) {
                for (int64_t k = 0; k < sd->n_bcs; ++k)
                    targets[k] = amp_y(sd->amps, sd->bc_amp[k], *timer, sd->bc_origin[k], step_id) * sd->bc_val[k];
                FX_HIP(hipMemcpyAsync(dbc.target.p, targets.data(), targets.size() * sizeof(double),
                                      hipMemcpyHostToDevice, s));
            }
            double rel_prev = -1.0;
            iters_prev.swap(iters_cur);
            iters_cur.clear();
            for (int it = 0; it < 15; ++it) {
                rep.nr_iterations++;
                fx::NewtonGraph::FrontArgs fa{nr_graph_ready, nlgeom, !host_solve, &opts, dmask.p, dbc.n_unique,
                                              dbc.ud_dof.p, dbc.ud_ptr.p, dbc.rec_order.p, dbc.target.p};
                FX_HIP(hipEventRecord(ev.e[0], s));
                ng.launch_front(fa);
                nr_graph_ready = use_graph;  // the first iteration runs plain launches, the rest the graph
                FX_HIP(hipEventRecord(ev.e[1], s));
                // the solve's iterations, queued without a status read: 15/16
                // of what the solve at this NR index of the previous
                // increment needed (else the last solve; 64 for the step's
                // first), none when this iteration is expected to converge
                // (a linear step after its first solve; NLGEOM when the
                // residual's last reduction carried on would pass 1e-3).
                // Launches past the stop rule return at once but are not
                // free (about a tenth of a wait each), so the queue stops a
                // little short and the chunked finish reads the rest.
                const bool expect_conv = it > 0 && (!nlgeom || (rel_prev > 0.0 && rep.last_rel_residual *
                                                                   (rep.last_rel_residual / rel_prev) < 1e-3));
                if (!host_solve && !expect_conv) {
                    const int64_t pred = (size_t)it < iters_prev.size() ? iters_prev[it]
                                         : !iters_cur.empty()            ? iters_cur.back()
                                                                         : 48;
                    fx::pcg_queue_fp64(m, G, pred - pred / 16);
                }
                ng.launch_post(!host_solve);
                FX_HIP(hipEventRecord(ev.e[2], s));
                const fx::NrReport &R = ng.wait();
                float ms_front = 0.f, ms_queued = 0.f;
                if (R.dev_code != 0)
                    throw Error{(fx_status)R.dev_code, element_error("Assembler::assemble_internal_force",
                                                                     R.dev_code, R.bad_element)};
                rel_prev = it > 0 ? rep.last_rel_residual : -1.0;
                rep.last_rel_residual = R.rel;
                FX_HIP(hipEventElapsedTime(&ms_front, ev.e[0], ev.e[1]));
                FX_HIP(hipEventElapsedTime(&ms_queued, ev.e[1], ev.e[2]));
                rep.seconds_assembly += 1e-3 * ms_front;
                if (R.skip) {
                    converged = true;
                    break;
                }
                memcpy(&m->last_max_diag, &R.max_diag_bits, sizeof(double));
                m->k_valid = true;
                fx_solve_report sr{};
                fx_status st = FX_OK;
                auto t_sol = clk::now();
                if (host_solve) {
                    st = sd->solve ? hook_solve(m, sd, &sr) : fx::pcg_solve(m, m->rhs.p, m->du.p, opts, &sr);
                    if (st == FX_OK)
                        fx::axpy_dev(m, n, 1.0, m->du.p, m->u.p);
                } else {
                    fx::PcgScalars h = R.pcg;
                    int code = h.status;
                    if (code == 0) {  // still running: the adaptive chunked finish, then du = x, u += du
                        code = fx::pcg_continue_fp64(m, opts.tol, opts.max_iter, G, &h);
                        if (code == 1) {
                            FX_HIP(hipMemcpyAsync(m->du.p, m->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
                            fx::axpy_dev(m, n, 1.0, m->du.p, m->u.p);
                        }
                    }
                    sr.iterations = fx::pcg_iterations_done(h);
                    sr.final_criterion = h.crit;
                    sr.spmv_group = G;
                    st = code == 2 ? FX_ERR_BREAKDOWN : (code == 3 ? FX_ERR_NOT_CONVERGED : FX_OK);
                    iters_cur.push_back(sr.iterations);
                }
                rep.seconds_solve += 1e-3 * ms_queued + since(t_sol);
                rep.linear_solves++;
                rep.pcg_iterations += sr.iterations;
                rep.outer_spmv += sr.outer_spmv;
                if (st != FX_OK) {
                    if (report) {
                        rep.host_syncs = fx::t_host_syncs - syncs0;
                        rep.host_syncs = fx::t_host_syncs - syncs0;
                    rep.seconds = since(t_start);
                        *report = rep;
                    }
                    throw Error{st, sd->solve ? "the step's linear solver (fx_step_def.solve) returned status " +
                                                    std::to_string((int)st)
                                              : solve_error(st, opts, sr)};
                }
            }
        } else
        for (int it = 0; it < 15; ++it) {
            rep.nr_iterations++;
            auto t_asm = clk::now();
            // 1. external force from amplitude-scaled loads, deck order
            set_fext(m, sd, *timer, step_id, touched, fext_h, tval, d_touched, d_tval);
            // 2. internal force at u_curr (coordinates move under NLGEOM);
            //    the thermal force is exactly zero (T never changes)
            internal_force_dev(m, nlgeom);
            // 3./4. residual and BC-masked norms
            fx::launch_residual(m, m->f_ext.p, m->f_int.p, m->rhs.p, n);
            fx::launch_masked_norms(m, m->rhs.p, m->f_ext.p, dmask.p, norms.p);
            double h2[2];
            FX_HIP(hipMemcpyAsync(h2, norms.p, sizeof h2, hipMemcpyDeviceToHost, s));
            fx::check_status(m, "Assembler::assemble_internal_force");  // synchronizes
            const double res_norm = std::sqrt(h2[0]), f_norm = std::sqrt(h2[1]);
            const double rel = f_norm > 1e-6 ? res_norm / f_norm : res_norm;
            rep.last_rel_residual = rel;
            if (rel < 1e-3) {
                converged = true;
                rep.seconds_assembly += since(t_asm);
                break;
            }
            // 5. tangent stiffness, 6. penalty boundary conditions
            double md = assemble_k(m, nlgeom);
            if (md > 0.0 && dbc.n_unique > 0) {
                for (int64_t k = 0; k < sd->n_bcs; ++k)
                    targets[k] = amp_y(sd->amps, sd->bc_amp[k], *timer, sd->bc_origin[k], step_id) * sd->bc_val[k];
                FX_HIP(hipMemcpyAsync(dbc.target.p, targets.data(), targets.size() * sizeof(double),
                                      hipMemcpyHostToDevice, s));
                fx::launch_penalty(m, dbc.n_unique, dbc.ud_dof.p, dbc.ud_ptr.p, dbc.target.p, dbc.rec_order.p,
                                   m->rhs.p, m->u.p);
            }
            FX_HIP(hipStreamSynchronize(s));
            ++fx::t_host_syncs;
            rep.seconds_assembly += since(t_asm);
            // 7. solve
            auto t_sol = clk::now();
            fx_solve_report sr{};
            fx_status st = sd->solve ? hook_solve(m, sd, &sr) : fx::pcg_solve(m, m->rhs.p, m->du.p, opts, &sr);
            rep.seconds_solve += since(t_sol);
            rep.linear_solves++;
            rep.pcg_iterations += sr.iterations;
            rep.outer_spmv += sr.outer_spmv;
            if (st != FX_OK) {
                if (report) {
                    rep.host_syncs = fx::t_host_syncs - syncs0;
                    rep.seconds = since(t_start);
                    *report = rep;
                }
                throw Error{st, sd->solve ? "the step's linear solver (fx_step_def.solve) returned status " +
                                                    std::to_string((int)st)
                                              : solve_error(st, opts, sr)};
            }
            // 8. u_curr += du
            fx::axpy_dev(m, n, 1.0, m->du.p, m->u.p);
        }
        if (converged) {
            current_time += dt;
            FX_HIP(hipMemcpyAsync(u_start.p, m->u.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
            if (sd->on_increment) {  // writer.write_increment(&inc_res, timer)
                FX_HIP(hipStreamSynchronize(s));
                if (sd->on_increment(sd->on_increment_user, step_id, (int32_t)n_inc, timer->global_time) != 0)
                    throw Error{FX_ERR_INVALID, "result writer failed at step " + std::to_string(step_id) +
                                                    ", increment " + std::to_string(n_inc)};
            }
        } else {
            rep.cutbacks++;
            dt /= 2.0;
            if (dt < sd->min_time_increment) {
                FX_HIP(hipMemcpyAsync(m->u.p, u_start.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
                FX_HIP(hipStreamSynchronize(s));
                if (report) {
                    rep.host_syncs = fx::t_host_syncs - syncs0;
                    rep.seconds = since(t_start);
                    *report = rep;
                }
                throw Error{FX_ERR_NOT_CONVERGED, "Minimum time increment reached. Convergence failed."};
            }
        }
    }
    // solution_state.displacements = u_step_start
    FX_HIP(hipMemcpyAsync(m->u.p, u_start.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    FX_HIP(hipStreamSynchronize(s));
    rep.host_syncs = fx::t_host_syncs - syncs0;
    rep.seconds = since(t_start);
    if (report)
        *report = rep;
    FX_API_END(m->ctx)
}

}  // extern "C"
