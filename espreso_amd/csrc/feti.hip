// Host driver of the MI355X FETI core and its C ABI (include/espreso_amd.h).
//
// Mirrors the reference's setup order (FETISolver::init, FETISolver.cpp:668-735):
// subdomain data -> lambda maps (cluster.cpp:33-220) -> coarse problem
// (itersolver.cpp:4931-5121) -> Dirichlet preconditioner and local dual
// operators (clustercpu.cpp:11-40, 167-333) -> projected PCG
// (itersolver.cpp:1406-1724) -> primal recovery (itersolver.cpp:372-431).
// Everything after the host preprocessing runs on one GPU; the host only
// issues launches and reads back the PCG residual norm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <type_traits>
#include <omp.h>
#include <map>
#include <mutex>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>

#include "dual.h"
#include "grid.h"
#include "internal.h"

using namespace esp;

namespace {

thread_local std::string g_last_error;

double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// std::vector storage without value-initialisation: the large host arrays of
// the set-up (R_b, GG^T, the lambda maps) are written in full by parallel
// loops, and zeroing them first would touch every page on one thread
template <class T>
struct NoInit : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <class U>
    NoInit(const NoInit<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using hvec = std::vector<T, NoInit<T>>;

struct SubHost {
    int n = 0, ni = 0, m = 0, kdim = 0;
    std::vector<int> rowptr, colidx;
    // column indices of a subdomain of at most 65536 DOFs, as 16-bit values
    // in pinned host memory (then colidx stays empty unless a 32-bit upload
    // needs it); col(k) reads whichever copy the subdomain has
    const uint16_t *col16 = nullptr;
    int64_t pin_pos = 0;  // position of col16 in the instance's pinned blocks laid end to end (16-bit units)
    int col(size_t k) const { return col16 ? (int)col16[k] : colidx[k]; }
    // the K values in pinned host memory (then vals stays empty), at vpin_pos
    // in the value blocks laid end to end (doubles); val_half: only the lower
    // triangle (the entries with column <= row, a prefix of each row, whose
    // columns are sorted), the device mirrors the rest (K is symmetric)
    double *val_pin = nullptr;
    int64_t vpin_pos = 0;
    bool val_half = false;
    // full values from the lower ones of this subdomain (rows sorted); false
    // when an entry's mirror is missing from the structure
    bool expand_half(const double *half, std::vector<double> &full) const {
        std::vector<int64_t> low(n + 1, 0);
        for (int r = 0; r < n; r++) {
            int c = 0;
            for (int q = rowptr[r]; q < rowptr[r + 1] && col(q) <= r; q++) c++;
            low[r + 1] = low[r] + c;
        }
        full.resize(rowptr[n]);
        for (int r = 0; r < n; r++)
            for (int q = rowptr[r]; q < rowptr[r + 1]; q++) {
                const int c = col(q);
                if (c <= r) {
                    full[q] = half[low[r] + (q - rowptr[r])];
                    continue;
                }
                int lo = rowptr[c], hi = rowptr[c + 1];
                while (lo < hi) {
                    const int mid = (lo + hi) / 2;
                    if (col(mid) < r) lo = mid + 1;
                    else hi = mid;
                }
                if (lo == rowptr[c + 1] || col(lo) != r) return false;
                full[q] = half[low[c] + (lo - rowptr[c])];
            }
        return true;
    }
    // the lower values (column <= row) of full into half, in CSR order
    void lower_values(const double *full, double *half) const {
        for (int r = 0; r < n; r++)
            for (int q = rowptr[r]; q < rowptr[r + 1]; q++)
                if (col(q) <= r) *half++ = full[q];
    }
    std::vector<double> vals, f, R;
    std::vector<int> lam, dof;
    std::vector<double> sign, scale;
    // derived
    std::vector<int> perm, iperm, bdofs, colstart;
    int band_tiles = 0, bw = 0;
    bool set = false;
    // hybrid FETI: cluster and B0 (corner) rows touching the subdomain
    int cluster = -1;
    std::vector<int> b0row, b0dof;
    std::vector<double> b0sign;
};

// symmetric eigen-decomposition (cyclic Jacobi) of the n x n column-major A;
// eigenvalues in w, eigenvectors in the columns of V
void jacobi_eigen(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[i + (size_t)i * n] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                tot += A[i + (size_t)j * n] * A[i + (size_t)j * n];
                if (i != j) off += A[i + (size_t)j * n] * A[i + (size_t)j * n];
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p + (size_t)q * n];
                if (apq == 0.0) continue;
                const double app = A[p + (size_t)p * n], aqq = A[q + (size_t)q * n];
                const double th = 0.5 * (aqq - app) / apq;
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = A[k + (size_t)p * n], akq = A[k + (size_t)q * n];
                    A[k + (size_t)p * n] = c * akp - sn * akq;
                    A[k + (size_t)q * n] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[p + (size_t)k * n], aqk = A[q + (size_t)k * n];
                    A[p + (size_t)k * n] = c * apk - sn * aqk;
                    A[q + (size_t)k * n] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
                    V[k + (size_t)p * n] = c * vkp - sn * vkq;
                    V[k + (size_t)q * n] = sn * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = A[i + (size_t)i * n];
}

// orthonormalise the k columns of V (n x k); false when they are dependent
bool mgs(int n, int k, double *V) {
    for (int pass = 0; pass < 2; pass++)
        for (int j = 0; j < k; j++) {
            double *v = V + (size_t)j * n;
            for (int i = 0; i < j; i++) {
                const double *u = V + (size_t)i * n;
                double d = 0;
                for (int r = 0; r < n; r++) d += u[r] * v[r];
                for (int r = 0; r < n; r++) v[r] -= d * u[r];
            }
            double nr = 0;
            for (int r = 0; r < n; r++) nr += v[r] * v[r];
            if (!(nr > 1e-24)) return false;
            nr = 1.0 / std::sqrt(nr);
            for (int r = 0; r < n; r++) v[r] *= nr;
        }
    return true;
}

}  // namespace

// The device's default memory pool keeps freed memory for the process
// (release threshold: no limit) instead of returning it to the driver at each
// synchronisation.
void keep_pooled_memory(int device) {
    static std::mutex mu;
    static std::vector<char> done;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)done.size() <= device) done.resize(device + 1, 0);
    if (done[device]) return;
    hipMemPool_t pool = nullptr;
    ESP_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = UINT64_MAX;
    ESP_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    done[device] = 1;
}

// The assembly workspaces (tens to hundreds of GB) are kept for the process
// once allocated and lent to the instances, largest first fit: memory that
// was in use before (by an earlier instance, or a block the stream-ordered
// pool gave back) is cleared by the driver at its next allocation, which
// takes seconds at this size.  ESPRESO_CACHE_WORKSPACE=0 frees them with
// their instance instead.
struct WorkspaceCache {
    struct Block {
        void *ptr;
        size_t bytes;
        bool busy;
    };
    std::mutex mu;
    std::map<int, std::vector<Block>> blocks;  // per device

    static bool keep() {
        const char *e = std::getenv("ESPRESO_CACHE_WORKSPACE");
        return !e || std::atoi(e) != 0;
    }
    // bytes of this device's blocks no instance uses
    size_t idle(int device) {
        std::lock_guard<std::mutex> lock(mu);
        size_t n = 0;
        for (const Block &b : blocks[device])
            if (!b.busy) n += b.bytes;
        return n;
    }
    // nullptr when the device has no room; an idle block is lent only if it
    // holds at most max_ratio times the request (0: any size)
    void *acquire(int device, size_t bytes, double max_ratio = 0.0) {
        std::lock_guard<std::mutex> lock(mu);
        std::vector<Block> &v = blocks[device];
        Block *best = nullptr;
        for (Block &b : v)
            if (!b.busy && b.bytes >= bytes && (max_ratio <= 0.0 || (double)b.bytes <= max_ratio * (double)bytes) &&
                (!best || b.bytes < best->bytes))
                best = &b;
        if (best) {
            best->busy = true;
            return best->ptr;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipErrorOutOfMemory) {
            // make room: give the idle blocks of the device back and retry
            (void)hipGetLastError();
            for (auto it = v.begin(); it != v.end();) {
                if (!it->busy) {
                    (void)hipFree(it->ptr);
                    it = v.erase(it);
                } else {
                    ++it;
                }
            }
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        v.push_back(Block{p, bytes, true});
        return p;
    }
    // frees the idle blocks of a device (room for allocations elsewhere)
    void release_idle(int device) {
        std::lock_guard<std::mutex> lock(mu);
        std::vector<Block> &v = blocks[device];
        for (auto it = v.begin(); it != v.end();) {
            if (!it->busy) {
                (void)hipFree(it->ptr);
                it = v.erase(it);
            } else {
                ++it;
            }
        }
    }
    void give_back(int device, void *p) {
        std::lock_guard<std::mutex> lock(mu);
        std::vector<Block> &v = blocks[device];
        for (auto it = v.begin(); it != v.end(); ++it)
            if (it->ptr == p) {
                if (keep()) {
                    it->busy = false;
                } else {
                    (void)hipFree(p);
                    v.erase(it);
                }
                return;
            }
    }
};
WorkspaceCache &workspace_cache() {
    static WorkspaceCache c;
    return c;
}

// Pinned host blocks that receive the 16-bit column indices of the K_i CSR
// at the handoff, so that they cross the link straight from there (no staging
// copy), kept for the process once allocated and lent to the instances:
// pinning a gigabyte costs tens of milliseconds.
struct PinnedPool {
    static constexpr size_t BLOCK = (size_t)64 << 20;
    std::mutex mu;
    std::vector<void *> idle;
    void *acquire() {
        {
            std::lock_guard<std::mutex> lock(mu);
            if (!idle.empty()) {
                void *p = idle.back();
                idle.pop_back();
                return p;
            }
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, BLOCK, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    void give_back(void *p) {
        std::lock_guard<std::mutex> lock(mu);
        idle.push_back(p);
    }
};
PinnedPool &pinned_pool() {
    static PinnedPool p;
    return p;
}

struct espreso_feti_s {
    espreso_feti_options opt{};
    int nsub = 0, nlam = -1, ncoarse = 0;
    std::vector<SubHost> sub;
    std::vector<double> c;
    hipStream_t st = nullptr;
    std::vector<void *> allocs;
    std::map<void *, size_t> alloc_size;  // bytes of each allocation
    size_t held_bytes = 0;                  // device memory the instance holds
    bool loaded = false, ready = false, solved = false;
    // grid bookkeeping
    int ndof = 0, nnodes = 0;
    std::vector<std::vector<int>> l2g;
    // device
    DualDev D;
    PrimalDev P;
    ClusterDev CL;
    bool hybrid = false;
    int ncl = 0;
    std::vector<std::vector<int>> cl_subs;          // subdomains of each cluster
    std::vector<int64_t> asubv;                     // kernel amplitude offsets per subdomain
    hvec<double> Rb_h, e0f_h;                       // host copies: R_b, -R_s^T f_s
    std::vector<int64_t> rboff_h;
    hvec<int> le_cnt_h, le_sub_h;
    hvec<int64_t> le_b_h;
    hvec<double> le_sign_h;
    std::vector<int> nq_h;
    std::vector<int64_t> qoff_h, fccoff_h;
    std::vector<int> qrow_h;                        // per B0 slot: global cluster row
    int64_t max_qe_h = 0;                           // entries in one B0 slot (1: corners)
    double *d_qj = nullptr, *d_Fcc = nullptr;       // (device) boundary index as double, corner blocks
    double *d_e0f = nullptr, *d_e0s = nullptr, *d_alpha_tot = nullptr;
    double *d_hf = nullptr, *d_ubf = nullptr, *d_fb = nullptr;  // boundary arrays
    double *d_x = nullptr, *d_r = nullptr, *d_w = nullptr, *d_y = nullptr, *d_z = nullptr, *d_p = nullptr,
           *d_Ap = nullptr, *d_b = nullptr, *d_tmp = nullptr, *d_r2 = nullptr, *d_partial = nullptr,
           *d_scal = nullptr;
    double *d_u = nullptr;    // per-subdomain primal (local order), sum n
    double *d_f = nullptr, *d_Qm = nullptr, *d_GGt = nullptr, *d_c = nullptr;  // resident inputs
    std::vector<int64_t> h_poff[3], h_boff;
    double *d_e = nullptr;    // -R^T f
    std::vector<int64_t> uoff;
    std::vector<double> lambda_h;
    std::vector<double> u_h;
    espreso_feti_stats last{};
    double fsc_bytes = 0, sd_bytes = 0, asm_flops = 0;
    // K values changed since the last set-up (espreso_feti_update_values): the
    // preconditioner values gathered from them are refreshed at the next set-up
    bool values_dirty = false;
    int64_t *d_kbb_pos = nullptr;                   // positions of the K_bb values in P.vals
    int64_t n_kbb = 0;
    // SUPER_DIRICHLET: K_bb - K_bi diag(K_ii)^-1 K_ib applied from the sparse blocks
    int64_t n_near = 0, n_sib = 0, n_sbi = 0;       // near-interior DOFs, K_ib and K_bi entries
    int64_t *d_sib_ptr = nullptr, *d_sib_pos = nullptr, *d_sbi_ptr = nullptr, *d_sbi_pos = nullptr,
            *d_sdiag_pos = nullptr, *d_sib_col = nullptr, *d_sbi_col = nullptr;
    double *d_sib_val = nullptr, *d_sbi_val = nullptr, *d_sdiag = nullptr, *d_snear = nullptr;
    // KKT report scratch (boundary vectors, lambda vector, per-subdomain sums,
    // their partial sums over row groups)
    double *d_kkt = nullptr;
    int kkt_max_n = 0;
    // pinned blocks holding the subdomains' 16-bit column indices and K values
    // (bump allocation; espreso_feti_set_subdomain may run in parallel)
    struct PinArena {
        std::vector<void *> blocks;
        size_t used = PinnedPool::BLOCK;  // bytes taken in the last block
        // bytes in the image of the blocks laid end to end
        size_t image_bytes() const { return blocks.size() * PinnedPool::BLOCK; }
    };
    std::mutex pin_mu;
    PinArena pin_idx, pin_val;
    // n values of type T in pinned memory of arena a, at pos (in T units) of
    // its image; nullptr when none is left or the request exceeds a block
    // (the caller keeps an ordinary host copy then)
    template <class T>
    T *pin_alloc(PinArena &a, size_t n, int64_t &pos) {
        const size_t bytes = (n * sizeof(T) + 63) & ~(size_t)63;
        if (bytes > PinnedPool::BLOCK) return nullptr;
        std::lock_guard<std::mutex> lock(pin_mu);
        if (a.used + bytes > PinnedPool::BLOCK) {
            void *b = pinned_pool().acquire();
            if (!b) return nullptr;
            a.blocks.push_back(b);
            a.used = 0;
        }
        T *p = (T *)((char *)a.blocks.back() + a.used);
        pos = (int64_t)((a.blocks.size() - 1) * PinnedPool::BLOCK + a.used) / (int64_t)sizeof(T);
        a.used += bytes;
        return p;
    }
    // the blocks of arena a into dst (an image of a.image_bytes()) on stream s
    void upload_arena(const PinArena &a, void *dst, hipStream_t s) {
        for (size_t b = 0; b < a.blocks.size(); b++)
            ESP_HIP(hipMemcpyAsync((char *)dst + b * PinnedPool::BLOCK, a.blocks[b],
                                   b + 1 < a.blocks.size() ? PinnedPool::BLOCK : a.used, hipMemcpyHostToDevice, s));
    }
    int measure = 0;                                // ESPRESO_MEASURE: phase times on stderr
    void note(const char *what, double t0) const {
        if (measure) std::fprintf(stderr, "espreso: %-28s %8.4f s\n", what, wall() - t0);
    }
    void gather_prec_values();
    void grid_stiffness();
    void kkt(const double *lambda, espreso_feti_stats &s);
    void truncate_primal();
    void check_launches(const char *where) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            throw Error(ESPRESO_ERR_HIP, std::string(where) + ": kernel launch failed: " + hipGetErrorString(e));
    }

    // TOTAL_FETI: GG^T is built and uploaded by a host thread while the GPU
    // assembles the local operators (it needs only R_b and the lambda maps);
    // coarse() joins it before reading d_GGt
    std::thread ggt_thread;
    std::exception_ptr ggt_err;
    hipStream_t ggt_st = nullptr;
    hipEvent_t ggt_ev = nullptr;
    void join_ggt() {
        if (ggt_thread.joinable()) ggt_thread.join();
        if (ggt_err) {
            std::exception_ptr e = ggt_err;
            ggt_err = nullptr;
            std::rethrow_exception(e);
        }
    }

    ~espreso_feti_s() {
        if (ggt_thread.joinable()) ggt_thread.join();
        if (ggt_ev) (void)hipEventDestroy(ggt_ev);
        if (ggt_st) (void)hipStreamDestroy(ggt_st);
        staging_release();
        if (!pin_idx.blocks.empty() || !pin_val.blocks.empty()) {
            if (st) (void)hipStreamSynchronize(st);  // the uploads from the blocks have finished
            for (PinArena *a : {&pin_idx, &pin_val}) {
                for (void *b : a->blocks) pinned_pool().give_back(b);
                a->blocks.clear();
            }
        }
        if (asm_pool || !cached_allocs.empty()) {
            if (st) (void)hipStreamSynchronize(st);
            if (asm_st) (void)hipStreamSynchronize(asm_st);
            for (Lane &l : lanes)
                if (l.s) (void)hipStreamSynchronize(l.s);
        }
        if (asm_pool) {
            workspace_cache().give_back(opt.device, asm_pool);
            asm_pool = nullptr;
        }
        for (void *p : cached_allocs) workspace_cache().give_back(opt.device, p);
        cached_allocs.clear();
        drop_graphs();
        for (Lane &l : lanes) {
            if (l.ev) (void)hipEventDestroy(l.ev);
            if (l.s) (void)hipStreamDestroy(l.s);
        }
        if (ev_main) (void)hipEventDestroy(ev_main);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        for (hipEvent_t e : asm_ev)
            if (e) (void)hipEventDestroy(e);
        if (asm_st) (void)hipStreamDestroy(asm_st);
        for (void *p : allocs) (void)hipFreeAsync(p, st);
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }
    // Device memory comes from the device's stream-ordered pool, which keeps
    // what instances free for the next ones of the process (like a caching
    // allocator): giving back and taking again the tens to hundreds of GB of
    // an instance costs seconds in the driver, reusing pooled memory nothing.
    // Allocations and frees are ordered on st; every other stream of the
    // instance starts behind an event of st.  Arrays of 256 MB and more come
    // from the process's block cache instead (a block holding at most twice
    // the request): the pool does not always find a free block of that size
    // again, and then the driver's fresh allocation costs a second or more.
    // nullptr when the device has no room (must_fit false), else an error
    static constexpr size_t CACHED_ALLOC_BYTES = (size_t)256 << 20;
    std::vector<void *> cached_allocs;
    void *alloc_bytes(size_t bytes, bool must_fit) {
        void *p = nullptr;
        if (bytes == 0) bytes = 8;
        if (bytes >= CACHED_ALLOC_BYTES && WorkspaceCache::keep()) {
            // blocks are used on st (and the streams joined to it) only after
            // this point; the cache holds only blocks whose users have finished
            const double ta = measure ? wall() : 0.0;
            p = workspace_cache().acquire(opt.device, bytes, 2.0);
            if (measure)
                std::fprintf(stderr, "espreso: cached block %.3f GB in %.4f s%s\n", 1e-9 * bytes, wall() - ta,
                             p ? "" : " (none: from the pool)");
            if (p) {
                cached_allocs.push_back(p);
                alloc_size[p] = bytes;
                held_bytes += bytes;
                return p;
            }
        }
        auto get = [&]() { return hipMallocAsync(&p, bytes, st); };
        const double ta = measure ? wall() : 0.0;
        hipError_t e = get();
        if (measure && bytes >= ((size_t)256 << 20))
            std::fprintf(stderr, "espreso: alloc %.3f GB in %.4f s%s\n", 1e-9 * bytes, wall() - ta,
                         e == hipErrorOutOfMemory ? " (out of memory, trimming the pool)" : "");
        if (e == hipErrorOutOfMemory) {
            // the pools' unused memory may sit in blocks that do not fit the
            // request, the block cache may hold idle blocks: give both back
            // to the driver and try once more
            (void)hipGetLastError();
            ESP_HIP(hipStreamSynchronize(st));
            hipMemPool_t pool = nullptr;
            ESP_HIP(hipDeviceGetDefaultMemPool(&pool, opt.device));
            ESP_HIP(hipMemPoolTrimTo(pool, 0));
            workspace_cache().release_idle(opt.device);
            e = get();
        }
        if (e == hipErrorOutOfMemory && !must_fit) {
            (void)hipGetLastError();
            return nullptr;
        }
        ESP_HIP(e);
        allocs.push_back(p);
        alloc_size[p] = bytes;
        held_bytes += bytes;
        return p;
    }
    template <class T>
    T *alloc(size_t n) { return (T *)alloc_bytes(n * sizeof(T), true); }
    // Uploads the concatenation of per-subdomain host pieces (piece(s) points
    // at off[s+1] - off[s] values) through two pinned staging buffers: the
    // host threads copy the pieces of one chunk into a buffer while the copy
    // engine drains the other.  Pageable copies of the same bytes run at less
    // than half the link rate, and a concatenated host copy first would add
    // two passes over them.
    static constexpr size_t STAGE_BYTES = (size_t)128 << 20;
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    double stage_wait_s = 0, stage_copy_s = 0;  // staged uploads: waits for the copy engine, host copies
    void staging_ready() {
        for (int b = 0; b < 2; b++) {
            if (!stage[b]) ESP_HIP(hipHostMalloc(&stage[b], STAGE_BYTES, hipHostMallocDefault));
            if (!stage_ev[b]) ESP_HIP(hipEventCreateWithFlags(&stage_ev[b], hipEventDisableTiming));
        }
    }
    void staging_release() {
        for (int b = 0; b < 2; b++) {
            if (stage_ev[b]) (void)hipEventSynchronize(stage_ev[b]);
            if (stage[b]) (void)hipHostFree(stage[b]);
            if (stage_ev[b]) (void)hipEventDestroy(stage_ev[b]);
            stage[b] = nullptr;
            stage_ev[b] = nullptr;
        }
    }
    // Copies the concatenation of the pieces into dst on stream s; threads
    // (0: the OpenMP default) fill the staging buffer, each with pieces of at
    // most 4 MB.  Pieces of another element type are converted to T on the way.
    template <class T, class Piece>
    void stage_pieces(T *dst, const std::vector<int64_t> &off, Piece piece, hipStream_t s, int threads = 0) {
        const int64_t total = off.back();
        if (total == 0) return;
        staging_ready();
        const int64_t chunk = (int64_t)(STAGE_BYTES / sizeof(T)), seg = (int64_t)(4 << 20) / (int64_t)sizeof(T);
        if (threads <= 0) threads = omp_get_max_threads();
        std::vector<std::array<int64_t, 3>> segs;  // piece, first value, end
        int buf = 0;
        for (int64_t c0 = 0; c0 < total; c0 += chunk, buf ^= 1) {
            const int64_t c1 = std::min(total, c0 + chunk);
            const int s0 = (int)(std::upper_bound(off.begin(), off.end(), c0) - off.begin()) - 1;
            const int s1 = (int)(std::lower_bound(off.begin(), off.end(), c1) - off.begin());
            segs.clear();
            for (int q = s0; q < s1; q++)
                for (int64_t a = std::max(off[q], c0), b = std::min(off[q + 1], c1); a < b; a += seg)
                    segs.push_back({q, a, std::min(b, a + seg)});
            const double w0 = wall();
            ESP_HIP(hipEventSynchronize(stage_ev[buf]));  // its previous chunk has left
            const double w1 = wall();
            T *sb = (T *)stage[buf];
#pragma omp parallel for schedule(dynamic) num_threads(threads)
            for (size_t g = 0; g < segs.size(); g++) {
                const int q = (int)segs[g][0];
                const int64_t a = segs[g][1], b = segs[g][2];
                const auto *src = piece(q) + (a - off[q]);
                if constexpr (std::is_same_v<std::remove_cv_t<std::remove_pointer_t<decltype(src)>>, T>) {
                    std::memcpy(sb + (a - c0), src, (size_t)(b - a) * sizeof(T));
                } else {
                    T *o = sb + (a - c0);  // narrowing copy (the caller checked the range)
                    for (int64_t e = 0; e < b - a; e++) o[e] = (T)src[e];
                }
            }
            stage_wait_s += w1 - w0;
            stage_copy_s += wall() - w1;
            ESP_HIP(hipMemcpyAsync(dst + c0, sb, (size_t)(c1 - c0) * sizeof(T), hipMemcpyHostToDevice, s));
            ESP_HIP(hipEventRecord(stage_ev[buf], s));
        }
    }
    template <class T, class Piece>
    T *upload_pieces(const std::vector<int64_t> &off, Piece piece) {
        T *dst = alloc<T>(off.back());
        stage_pieces(dst, off, piece, st);
        return dst;
    }
    // vectors of more than 8 MB go through the pinned staging buffers
    template <class T, class A>
    T *upload(const std::vector<T, A> &v) {
        if (v.size() * sizeof(T) > ((size_t)8 << 20)) {
            const std::vector<int64_t> off = {0, (int64_t)v.size()};
            return upload_pieces<T>(off, [&](int) { return v.data(); });
        }
        T *p = alloc<T>(v.size());
        if (!v.empty()) ESP_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
        return p;
    }
    void release(void *p) {
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) {
            (void)hipFreeAsync(p, st);
            allocs.erase(it);
            held_bytes -= alloc_size[p];
            alloc_size.erase(p);
            return;
        }
        auto ic = std::find(cached_allocs.begin(), cached_allocs.end(), p);
        if (ic != cached_allocs.end()) {
            // its users run on st: they finish before the block is lent again
            ESP_HIP(hipStreamSynchronize(st));
            workspace_cache().give_back(opt.device, p);
            cached_allocs.erase(ic);
            held_bytes -= alloc_size[p];
            alloc_size.erase(p);
        }
    }
    void sync() { ESP_HIP(hipStreamSynchronize(st)); }

    void load();
    void gpu_setup();
    void assemble(const double *d_f, const double *d_Q, const std::vector<int64_t> *poff,
                  const std::vector<int64_t> &boff);
    void coarse();
    void coarse_layout_subdomains();
    void load_clusters();
    double *d_alpha_c = nullptr;    // coarse amplitudes
    template <class T>
    T *replace_upload(T *old, const std::vector<T> &v) {
        if (old) release(old);
        return upload(v);
    }
    void cluster_setup();
    void pcg_step(double *x, double *r, int prec, bool first);
    void pcg_step_graph(double *x, double *r, int prec);
    void drop_graphs();
    struct StepGraph {
        double *x, *r;
        int prec;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> step_graphs;
    double *asm_pool = nullptr;  // assembly workspace, reused by later set-ups
    size_t asm_pool_n = 0;
    std::vector<int> wave_plan;  // end subdomain of each assembly wave
    bool use_graphs = true;
    void compute_b();
    void compute_b_into(const double *hf, double *ubf, const double *e0f, double *b);
    void rhs_host(const double *f, hvec<double> &e0f, std::vector<double> &e) const;
    void resolve_rhs();
    // block re-solve of several load cases (groups of up to 16)
    struct BlockBuffers {
        int K = 0;
        double *x = nullptr, *r = nullptr, *w = nullptr, *z = nullptr, *y = nullptr, *p = nullptr, *Ap = nullptr,
               *b = nullptr, *e = nullptr, *alpha_c = nullptr, *alpha_tot = nullptr, *e0f = nullptr, *hf = nullptr,
               *ubf = nullptr, *xbv = nullptr, *yb = nullptr, *f = nullptr, *yf = nullptr, *u = nullptr,
               *scal = nullptr, *work = nullptr, *e0s = nullptr, *lam0 = nullptr, *alphaK = nullptr;
    } BB;
    SymvBlock SB;
    int64_t nyf = 0;  // interior DOFs (length of P.yf)
    int64_t n_dsub = 1, n_lambda0 = 1, n_e0s = 1;  // scratch sizes (per-case lanes)
    // per-case work of a block runs on NLANE streams (default priority), each
    // with its own copies of the scratch the projector, the cluster correction
    // and the dot products use
    static constexpr int NLANE = 4;
    struct Lane {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        double *dsub = nullptr, *cd = nullptr, *ce = nullptr, *lambda0 = nullptr, *alphaK = nullptr,
               *e0s = nullptr, *partial = nullptr;
        DualDev D;
        ClusterDev C;
    } lanes[NLANE];
    hipEvent_t ev_main = nullptr;
    // assembly kernels that only read what an earlier step wrote run on asm_st
    // beside the main stream's work: fork (asm_st behind st) and join (st
    // behind asm_st)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void asm_fork() {
        if (!asm_st) ESP_HIP(hipStreamCreate(&asm_st));
        if (!ev_fork) ESP_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        ESP_HIP(hipEventRecord(ev_fork, st));
        ESP_HIP(hipStreamWaitEvent(asm_st, ev_fork, 0));
    }
    void asm_join() {
        if (!ev_join) ESP_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        ESP_HIP(hipEventRecord(ev_join, asm_st));
        ESP_HIP(hipStreamWaitEvent(st, ev_join, 0));
    }
    // assembly: forward substitution on a second stream (default priority), one
    // event per Cholesky step plus the join
    hipStream_t asm_st = nullptr;
    std::vector<hipEvent_t> asm_ev;
    void factor_solve(const DenseDesc *desc, int nb, int which, int nt, int band, int ntc);
    bool lanes_ready = false;
    void lanes_setup();
    void fork_lanes();
    void join_lanes();
    void block_setup(int K);
    void apply_F_block(int K, const double *x, double *y);
    void cluster_block(int K, const double *x);
    void step_block(int K, const std::vector<char> &active, bool first);
    // iterations with every case of a group active replay a HIP graph of
    // step_block (the lanes are forked and joined inside the graph)
    struct BlockGraph {
        int K;
        hipGraphExec_t exec;
    };
    std::vector<BlockGraph> block_graphs;
    void step_block_graph(int K);
    bool resolve_block(int nrhs, const double *f, double *u, int32_t *iters, double *norms,
                       espreso_feti_stats *stats);
    void resolve_group(int K, const double *f, double *u, int32_t *iters, double *norms, espreso_feti_stats &tot);
    std::vector<double> Ns_h;                        // hybrid: N_s per subdomain (kdim x kN)
    std::vector<int64_t> noff_h;
    std::vector<int> ccoff_h, cblk_h;                // coarse block offset, block of each subdomain
    int gridNE[3] = {0, 0, 0};                       // structured grid front end
    std::vector<double> gridKe0;
    bool grid_dev = false;                           // K_i values assembled on the device
    std::vector<int> grid_box;                       // 6 per subdomain
    std::vector<double> grid_escale;                 // element material scales
    std::vector<int64_t> h_nzoff;                    // CSR value offsets per subdomain
    void invert_spd_batched(double *pool, const std::vector<int64_t> &off, const std::vector<int> &n,
                            const char *what);
    std::vector<int> n0_h, kt_h;
    std::vector<int64_t> f0off_h, saoff_h;
    void build_ggt(const double *Rt, const std::vector<int64_t> &rtoff, const std::vector<int> &ck,
                   const std::vector<int> &cblk, const std::vector<int> &coff, hvec<double> &GGt);
    void build_items(int prec, int64_t target, std::vector<SymvItem> &items, std::vector<int> &item0,
                     std::vector<int> &nit, int64_t &plen);
    double tile_flops(const SubHost &h) const;
    // SYMV work-item size (matrix elements per workgroup) per storage
    // precision (FP64, FP32, BF16); 0 for a precision nothing is stored in;
    // chosen by tune_symv
    int64_t symv_target[3] = {0, 0, 0};
    bool symv_tuned = false;
    int max_m_h = 0;
    bool prec_used(int p) const { return p == 0 || p == opt.sc_precision || p == opt.sd_precision; }
    void upload_items(int prec, int64_t target);
    void tune_symv();
    // dual operators on device vectors
    void apply_F(const double *x, double *y, int prec);
    void apply_prec(const double *x, double *y);
    void solve(espreso_feti_stats *st_out);
    void pcg(const double *b, double *x, double *r, int prec, int &iters, double &norm, double &tol, bool quiet_tol,
             double tol_abs, double *norm_b = nullptr);
};

// ----------------------------------------------------------------- setup ----
void espreso_feti_s::load() {
    ESP_CHECK(nlam >= 0, ESPRESO_ERR_STATE, "espreso_feti_set_dual not called");
    for (int s = 0; s < nsub; s++) ESP_CHECK(sub[s].set, ESPRESO_ERR_STATE, "subdomain " + std::to_string(s) + " not set");
    double t0 = wall();
    if (const char *e = std::getenv("ESPRESO_MEASURE")) measure = std::atoi(e);
    const int prec = opt.sc_precision;
    ESP_CHECK(prec >= 0 && prec <= 2, ESPRESO_ERR_ARG, "sc_precision must be 0, 1 or 2");
    ESP_CHECK(opt.preconditioner >= 0 && opt.preconditioner <= 4, ESPRESO_ERR_ARG,
              "preconditioner must be NONE, LUMPED, WEIGHT_FUNCTION, DIRICHLET or SUPER_DIRICHLET (0..4)");

    // The K_i CSR arrays need no preprocessing: their upload (16-bit column
    // indices and K values straight from the pinned blocks they were handed
    // over into, the rest staged by eight copy threads) and their placement on
    // the device run on a stream of their own, queued by a host thread beside
    // the host work below; the main stream waits for it only before the first
    // kernel that reads them, so the other uploads overlap the transfer
    std::vector<int64_t> nz_off(nsub + 1, 0), rp_off(nsub + 1, 0);
    for (int s = 0; s < nsub; s++) {
        nz_off[s + 1] = nz_off[s] + sub[s].rowptr[sub[s].n];
        rp_off[s + 1] = rp_off[s] + sub[s].n + 1;
    }
    P.rowptr = alloc<int>(rp_off[nsub]);
    P.colidx = alloc<int>(nz_off[nsub]);
    P.vals = alloc<double>(nz_off[nsub]);
    // column indices of subdomains with at most 65536 DOFs cross the link as
    // 16-bit values (half the bytes of the largest input) and are widened on
    // the device
    int max_n_in = 0;
    for (int s = 0; s < nsub; s++) max_n_in = std::max(max_n_in, sub[s].n);
    int n_pinned = 0;
    for (int s = 0; s < nsub; s++) n_pinned += sub[s].col16 != nullptr;
    // every subdomain's indices pinned at the handoff: the pinned blocks go up
    // whole (a few large copies; one copy per subdomain costs more in calls
    // than in bytes) into an image of the blocks, and a gather widens each
    // subdomain's run into place on the device
    const bool all_pinned = max_n_in <= 65536 && n_pinned == nsub && nsub > 0;
    if (measure) {
        int n_half = 0;
        for (int s = 0; s < nsub; s++) n_half += sub[s].val_half;
        std::fprintf(stderr,
                     "espreso: load: pinned at the handoff: 16-bit column indices of %d, K values of %d (%d as lower "
                     "triangles) of %d subdomains\n",
                     n_pinned, (int)std::count_if(sub.begin(), sub.end(), [](const SubHost &h) { return h.val_pin; }),
                     n_half, nsub);
    }
    uint16_t *pin_img = all_pinned ? (uint16_t *)alloc_bytes(pin_idx.image_bytes(), true) : nullptr;
    int n_vpinned = 0;
    for (int s = 0; s < nsub; s++) n_vpinned += sub[s].val_pin != nullptr;
    const bool vals_pinned = !grid_dev && n_vpinned == nsub && nsub > 0;
    double *vpin_img = vals_pinned ? (double *)alloc_bytes(pin_val.image_bytes(), true) : nullptr;
    uint16_t *colidx16 = max_n_in <= 65536 && !all_pinned ? alloc<uint16_t>(nz_off[nsub]) : nullptr;
    // where each subdomain's run sits in the images, and the largest run
    std::vector<int64_t> pin_pos_h(nsub), vpin_pos_h(nsub);
    int64_t max_nnz = 0;
    for (int s = 0; s < nsub; s++) {
        pin_pos_h[s] = sub[s].pin_pos;
        vpin_pos_h[s] = sub[s].vpin_pos;
        max_nnz = std::max(max_nnz, nz_off[s + 1] - nz_off[s]);
    }
    int64_t *d_pos = (pin_img || vpin_img) ? alloc<int64_t>(3 * (size_t)nsub + 1) : nullptr;
    // values handed over as lower triangles: mirrored on the device
    std::vector<int> half_h(nsub);
    bool any_half = false;
    for (int s = 0; s < nsub; s++) {
        half_h[s] = sub[s].val_half;
        any_half = any_half || sub[s].val_half;
    }
    any_half = any_half && vals_pinned;
    int *d_half = any_half ? alloc<int>(nsub + 1) : nullptr;  // [nsub]: the missing-mirror flag
    int *d_low = any_half ? alloc<int>(rp_off[nsub]) : nullptr;
    int64_t *d_rpo = any_half ? alloc<int64_t>(nsub + 1) : nullptr;
    std::atomic<int> sym_err{0};
    // the upload stream starts after the allocations above (stream-ordered on st)
    struct UpStream {
        hipStream_t s = nullptr;
        hipEvent_t ready = nullptr, done = nullptr;
        ~UpStream() {
            if (s) (void)hipStreamSynchronize(s);
            if (ready) (void)hipEventDestroy(ready);
            if (done) (void)hipEventDestroy(done);
            if (s) (void)hipStreamDestroy(s);
        }
    } up;
    ESP_HIP(hipStreamCreate(&up.s));
    ESP_HIP(hipEventCreateWithFlags(&up.ready, hipEventDisableTiming));
    ESP_HIP(hipEventCreateWithFlags(&up.done, hipEventDisableTiming));
    ESP_HIP(hipEventRecord(up.ready, st));
    std::exception_ptr csr_err;
    double csr_t[3] = {0, 0, 0};  // staging buffers ready, row pointers queued, last chunk queued (measure mode)
    struct Joined {
        std::thread t;
        ~Joined() {
            if (t.joinable()) t.join();
        }
    } csr_up;
    // 16-bit indices already pinned for every subdomain go up as they are
    csr_up.t = std::thread([&] {
        try {
            ESP_HIP(hipSetDevice(opt.device));
            ESP_HIP(hipStreamWaitEvent(up.s, up.ready, 0));
            staging_ready();
            csr_t[0] = wall() - t0;
            stage_pieces(P.rowptr, rp_off, [&](int s) { return sub[s].rowptr.data(); }, up.s, 8);
            csr_t[2] = wall() - t0;
            if (d_pos) {
                ESP_HIP(hipMemcpyAsync(d_pos, pin_pos_h.data(), nsub * sizeof(int64_t), hipMemcpyHostToDevice, up.s));
                ESP_HIP(hipMemcpyAsync(d_pos + nsub, vpin_pos_h.data(), nsub * sizeof(int64_t), hipMemcpyHostToDevice,
                                       up.s));
                ESP_HIP(hipMemcpyAsync(d_pos + 2 * nsub, nz_off.data(), (nsub + 1) * sizeof(int64_t),
                                       hipMemcpyHostToDevice, up.s));
            }
            if (all_pinned) {
                upload_arena(pin_idx, pin_img, up.s);
                launch_widen_runs(pin_img, d_pos, d_pos + 2 * nsub, nsub, max_nnz, P.colidx, up.s);
            } else {
                // mixed: the staged path reads 32-bit indices of every subdomain
#pragma omp parallel for schedule(dynamic) num_threads(8)
                for (int s = 0; s < nsub; s++) {
                    SubHost &h = sub[s];
                    if (h.col16 && h.colidx.empty())
                        h.colidx.assign(h.col16, h.col16 + h.rowptr[h.n]);
                }
                if (colidx16) {
                    stage_pieces(colidx16, nz_off, [&](int s) { return sub[s].colidx.data(); }, up.s, 8);
                    launch_widen_u16(colidx16, P.colidx, nz_off[nsub], up.s);
                } else {
                    stage_pieces(P.colidx, nz_off, [&](int s) { return sub[s].colidx.data(); }, up.s, 8);
                }
            }
            // the structured grid assembles the values on the device
            if (vals_pinned) {
                upload_arena(pin_val, vpin_img, up.s);
                if (any_half) {
                    half_h.push_back(0);
                    ESP_HIP(hipMemcpyAsync(d_half, half_h.data(), (nsub + 1) * sizeof(int), hipMemcpyHostToDevice,
                                           up.s));
                    ESP_HIP(hipMemcpyAsync(d_rpo, rp_off.data(), (nsub + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                                           up.s));
                    launch_expand_sym(vpin_img, d_pos + nsub, d_pos + 2 * nsub, P.rowptr, d_rpo, P.colidx, d_half,
                                      d_low, nsub, max_n_in, P.vals, d_half + nsub, up.s);
                } else {
                    launch_copy_runs(vpin_img, d_pos + nsub, d_pos + 2 * nsub, nsub, max_nnz, P.vals, up.s);
                }
            } else if (!grid_dev) {
#pragma omp parallel for schedule(dynamic) num_threads(8)
                for (int s = 0; s < nsub; s++) {
                    SubHost &h = sub[s];
                    if (h.val_pin && h.vals.empty()) {
                        if (!h.val_half) h.vals.assign(h.val_pin, h.val_pin + h.rowptr[h.n]);
                        else if (!h.expand_half(h.val_pin, h.vals)) sym_err = 1;
                    }
                }
                stage_pieces(P.vals, nz_off, [&](int s) { return sub[s].vals.data(); }, up.s, 8);
            }
            ESP_HIP(hipEventRecord(up.done, up.s));
            csr_t[1] = wall() - t0;
        } catch (...) {
            csr_err = std::current_exception();
        }
    });

    // ---- host preprocessing: boundary sets, permutations, bands ----
    std::vector<int64_t> boff(nsub + 1, 0), coffv(nsub + 1, 0);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; s++) {
        SubHost &h = sub[s];
        std::vector<int> b(h.dof);
        b.insert(b.end(), h.b0dof.begin(), h.b0dof.end());
        std::sort(b.begin(), b.end());
        b.erase(std::unique(b.begin(), b.end()), b.end());
        h.bdofs = b;
        h.m = (int)b.size();
        h.ni = h.n - h.m;
        h.iperm.assign(h.n, -1);
        for (int j = 0; j < h.m; j++) h.iperm[b[j]] = h.ni + j;
        h.perm.assign(h.n, 0);
        int q = 0;
        for (int i = 0; i < h.n; i++)
            if (h.iperm[i] < 0) { h.iperm[i] = q; h.perm[q++] = i; }
        for (int j = 0; j < h.m; j++) h.perm[h.ni + j] = b[j];
        // band of K_ii and first coupled interior row of each K_ib column
        int bw = 0, bt = 0;
        std::vector<int> first_int(h.m + 1, h.ni);
        first_int[h.m] = 0;  // the f column is dense
        for (int r = 0; r < h.ni; r++) {
            const int lr = h.perm[r];
            int fr = r;
            for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) {
                const int pc = h.iperm[h.col(k)];
                if (pc < h.ni) fr = std::min(fr, pc);
                else first_int[pc - h.ni] = std::min(first_int[pc - h.ni], r);
            }
            bw = std::max(bw, r - fr);
            bt = std::max(bt, r / TILE - fr / TILE);
        }
        h.bw = bw;
        h.band_tiles = std::max(bt, 1);
        const int ntc = (h.m + 1 + TILE - 1) / TILE;
        h.colstart.assign(ntc, 0);
        for (int J = 0; J < ntc; J++) {
            int mn = h.ni;
            for (int j = J * TILE; j < std::min(h.m + 1, (J + 1) * TILE); j++) mn = std::min(mn, first_int[j]);
            h.colstart[J] = mn / TILE;
        }
    }
    int max_m = 0, max_m_sub = 0;
    for (int s = 0; s < nsub; s++) {
        boff[s + 1] = boff[s] + sub[s].m;
        coffv[s + 1] = coffv[s] + sub[s].kdim;
        if (sub[s].m > max_m) { max_m = sub[s].m; max_m_sub = s; }
    }
    note("load: boundary sets, bands", t0);
    {
        // the SYMV keeps x and the row/column sums of a work item's columns in
        // LDS (symv_lds_bytes, plus the workgroup's max|x| slots)
        int maxsh = 0;
        ESP_HIP(hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, opt.device));
        for (int p = 0; p < 3; p++) {
            if (!prec_used(p)) continue;
            const int64_t need = symv_lds_bytes(max_m, p) + 4 * 8;
            ESP_CHECK(need <= maxsh, ESPRESO_ERR_ARG,
                      "subdomain " + std::to_string(max_m_sub) + " has " + std::to_string(max_m) +
                          " interface DOFs: its dual-operator workgroup would need " + std::to_string(need) +
                          " bytes of LDS (device limit " + std::to_string(maxsh) +
                          "); decompose into more subdomains");
        }
    }
    asubv = coffv;
    const int64_t nb = boff[nsub];

    // boundary -> lambda CSR and lambda -> entries.  Every boundary index
    // belongs to one subdomain, so the subdomains fill their ranges of the
    // CSR in parallel.  A lambda's entries keep the order of (subdomain,
    // position in its lambda list): the entry with the smaller key is
    // entry 0, found with an atomic minimum over the keys.
    std::vector<int64_t> lam_base(nsub + 1, 0);
    for (int s = 0; s < nsub; s++) lam_base[s + 1] = lam_base[s] + (int64_t)sub[s].lam.size();
    const int64_t nle = lam_base[nsub];
    hvec<int64_t> bl_ptr(nb + 1);
    hvec<int> bl_lam(nle);
    hvec<double> bl_sign(nle), bl_ss(nle);
    hvec<int> le_cnt(nlam), le_sub(2 * (size_t)nlam);
    hvec<int64_t> le_b(2 * (size_t)nlam), le_key(nlam);
    hvec<double> le_sign(2 * (size_t)nlam), le_ss(2 * (size_t)nlam);
    bl_ptr[0] = 0;
#pragma omp parallel for schedule(static)
    for (int l = 0; l < nlam; l++) {
        le_cnt[l] = 0;
        le_key[l] = INT64_MAX;
        le_sub[2 * (size_t)l] = le_sub[2 * (size_t)l + 1] = 0;
        le_b[2 * (size_t)l] = le_b[2 * (size_t)l + 1] = 0;
        le_sign[2 * (size_t)l] = le_sign[2 * (size_t)l + 1] = 0.0;
        le_ss[2 * (size_t)l] = le_ss[2 * (size_t)l + 1] = 0.0;
    }
    std::atomic<bool> lam_range_ok{true};
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        int64_t *cnt = bl_ptr.data() + boff[s] + 1;  // counts, then ends of the rows
        std::fill(cnt, cnt + h.m, 0);
        for (size_t q = 0; q < h.lam.size(); q++) {
            cnt[h.iperm[h.dof[q]] - h.ni]++;
            const int l = h.lam[q];
            if (l < 0 || l >= nlam) {
                lam_range_ok = false;
                continue;
            }
            __atomic_fetch_add(&le_cnt[l], 1, __ATOMIC_RELAXED);
            const int64_t key = ((int64_t)s << 32) | (int64_t)q;
            int64_t cur = __atomic_load_n(&le_key[l], __ATOMIC_RELAXED);
            while (key < cur && !__atomic_compare_exchange_n(&le_key[l], &cur, key, true, __ATOMIC_RELAXED,
                                                              __ATOMIC_RELAXED)) {
            }
        }
        int64_t acc = lam_base[s];
        for (int j = 0; j < h.m; j++) {
            const int64_t c = cnt[j];
            cnt[j] = acc;  // start of row j for now
            acc += c;
        }
    }
    ESP_CHECK(lam_range_ok, ESPRESO_ERR_ARG, "lambda id out of range");
    for (int l = 0; l < nlam; l++) {
        ESP_CHECK(le_cnt[l] > 0, ESPRESO_ERR_ARG, "lambda row without entries");
        ESP_CHECK(le_cnt[l] <= 2, ESPRESO_ERR_ARG, "a lambda row may touch at most two subdomain DOFs");
    }
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        int64_t *cur = bl_ptr.data() + boff[s] + 1;  // row starts; each ends as the row's end
        for (size_t q = 0; q < h.lam.size(); q++) {
            const int l = h.lam[q];
            const int64_t bj = boff[s] + h.iperm[h.dof[q]] - h.ni;
            const int64_t k = cur[bj - boff[s]]++;
            bl_lam[k] = l;
            bl_sign[k] = h.sign[q];
            bl_ss[k] = h.sign[q] * h.scale[q];
            const int e = le_key[l] == (((int64_t)s << 32) | (int64_t)q) ? 0 : 1;
            le_b[2 * (size_t)l + e] = bj;
            le_sub[2 * (size_t)l + e] = s;
            le_sign[2 * (size_t)l + e] = h.sign[q];
            le_ss[2 * (size_t)l + e] = h.sign[q] * h.scale[q];
        }
    }
    note("load: lambda maps", t0);

    // R_b, Q = orth(R_b), fb, K_bb (boundary CSR), e = -R^T f, GG^T
    std::vector<int64_t> rboff(nsub + 1, 0);
    for (int s = 0; s < nsub; s++) rboff[s + 1] = rboff[s] + (int64_t)sub[s].m * sub[s].kdim;
    hvec<double> Rb(rboff[nsub]), Q(rboff[nsub]), fb(nb), e(coffv[nsub]);
    std::vector<int64_t> kbb_cnt(nsub);
    std::vector<char> rank_ok(nsub, 1);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        for (int a = 0; a < h.kdim; a++)
            for (int j = 0; j < h.m; j++) Rb[rboff[s] + j + (int64_t)a * h.m] = h.R[h.bdofs[j] + (size_t)a * h.n];
        std::copy(Rb.begin() + rboff[s], Rb.begin() + rboff[s + 1], Q.begin() + rboff[s]);
        rank_ok[s] = mgs(h.m, h.kdim, Q.data() + rboff[s]);
        for (int j = 0; j < h.m; j++) fb[boff[s] + j] = h.f[h.bdofs[j]];
        for (int a = 0; a < h.kdim; a++) {
            double acc = 0;
            for (int i = 0; i < h.n; i++) acc += h.R[i + (size_t)a * h.n] * h.f[i];
            e[coffv[s] + a] = -acc;
        }
        int64_t cnt = 0;
        for (int j = 0; j < h.m; j++) {
            const int lr = h.bdofs[j];
            for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) cnt += h.iperm[h.col(k)] >= h.ni;
        }
        kbb_cnt[s] = cnt;
    }
    for (int s = 0; s < nsub; s++)
        ESP_CHECK(rank_ok[s], ESPRESO_ERR_NUMERIC,
                  "kernel restricted to the boundary of subdomain " + std::to_string(s) + " is rank deficient");
    note("load: R_b, Q, e", t0);
    // K_bb in boundary indices: columns and positions of the values in the
    // concatenated CSR value array (gathered on the device once K is there).
    // SUPER_DIRICHLET also needs K_ib for the interior DOFs next to the
    // boundary ("near" DOFs, global index over the subdomains), their
    // diagonal, and K_bi by boundary row.
    h_nzoff.assign(nsub + 1, 0);
    for (int s = 0; s < nsub; s++) h_nzoff[s + 1] = h_nzoff[s] + sub[s].rowptr[sub[s].n];
    const bool superdir = opt.preconditioner == ESPRESO_PREC_SUPER_DIRICHLET;
    // K_bb is applied by the lumped and super-Dirichlet preconditioners only
    const bool need_kbb = superdir || opt.preconditioner == ESPRESO_PREC_LUMPED;
    std::vector<int64_t> kbb_off(nsub + 1, 0);
    for (int s = 0; s < nsub; s++) kbb_off[s + 1] = kbb_off[s] + (need_kbb ? kbb_cnt[s] : 0);
    std::vector<int64_t> kbb_ptr(nb + 1, 0), kbb_col(kbb_off[nsub]), kbb_pos(kbb_off[nsub]);
    std::vector<std::vector<int>> near_of(superdir ? nsub : 0);  // per subdomain: near DOFs (internal index)
    std::vector<int64_t> near_off(nsub + 1, 0), sib_off(nsub + 1, 0), sbi_off(nsub + 1, 0);
    if (superdir) {
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < nsub; s++) {
            const SubHost &h = sub[s];
            int64_t cib = 0, cbi = 0;
            for (int r = 0; r < h.ni; r++) {
                const int lr = h.perm[r];
                int nbc = 0;
                for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) nbc += h.iperm[h.col(k)] >= h.ni;
                if (nbc) { near_of[s].push_back(r); cib += nbc; }
            }
            for (int j = 0; j < h.m; j++) {
                const int lr = h.bdofs[j];
                for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) cbi += h.iperm[h.col(k)] < h.ni;
            }
            sib_off[s + 1] = cib;
            sbi_off[s + 1] = cbi;
        }
        for (int s = 0; s < nsub; s++) {
            near_off[s + 1] = near_off[s] + (int64_t)near_of[s].size();
            sib_off[s + 1] += sib_off[s];
            sbi_off[s + 1] += sbi_off[s];
        }
    }
    std::vector<int64_t> sib_ptr(superdir ? near_off[nsub] + 1 : 0, 0), sib_col(sib_off[nsub]), sib_pos(sib_off[nsub]),
        sdiag_pos(near_off[nsub]), sbi_ptr(superdir ? nb + 1 : 0, 0), sbi_col(sbi_off[nsub]), sbi_pos(sbi_off[nsub]);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        if (!need_kbb) continue;
        int64_t q = kbb_off[s];
        for (int j = 0; j < h.m; j++) {
            const int lr = h.bdofs[j];
            for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) {
                const int pc = h.iperm[h.col(k)];
                if (pc >= h.ni) {
                    kbb_col[q] = boff[s] + pc - h.ni;
                    kbb_pos[q] = h_nzoff[s] + k;
                    q++;
                }
            }
            kbb_ptr[boff[s] + j + 1] = q;
        }
        if (!superdir) continue;
        std::vector<int> near_idx(h.ni, -1);
        for (size_t a = 0; a < near_of[s].size(); a++) near_idx[near_of[s][a]] = (int)a;
        int64_t qi = sib_off[s];
        for (size_t a = 0; a < near_of[s].size(); a++) {
            const int r = near_of[s][a], lr = h.perm[r];
            for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) {
                const int pc = h.iperm[h.col(k)];
                if (pc >= h.ni) {
                    sib_col[qi] = boff[s] + pc - h.ni;
                    sib_pos[qi] = h_nzoff[s] + k;
                    qi++;
                } else if (pc == r) {
                    sdiag_pos[near_off[s] + a] = h_nzoff[s] + k;
                }
            }
            sib_ptr[near_off[s] + a + 1] = qi;
        }
        int64_t qb = sbi_off[s];
        for (int j = 0; j < h.m; j++) {
            const int lr = h.bdofs[j];
            for (int k = h.rowptr[lr]; k < h.rowptr[lr + 1]; k++) {
                const int pc = h.iperm[h.col(k)];
                if (pc < h.ni) {
                    sbi_col[qb] = near_off[s] + near_idx[pc];
                    sbi_pos[qb] = h_nzoff[s] + k;
                    qb++;
                }
            }
            sbi_ptr[boff[s] + j + 1] = qb;
        }
    }
    if (superdir) {
        for (int s = 0; s < nsub; s++)
            for (size_t a = 0; a < near_of[s].size(); a++)
                ESP_CHECK(sdiag_pos[near_off[s] + a] >= h_nzoff[s], ESPRESO_ERR_ARG,
                          "SUPER_DIRICHLET: K_" + std::to_string(s) + " lacks a diagonal entry");
        n_near = near_off[nsub];
        n_sib = sib_off[nsub];
        n_sbi = sbi_off[nsub];
    }
    note("load: K_bb (and K_ib) maps", t0);
    rboff_h = rboff;
    hybrid = opt.method == 1;
    if (hybrid) {
        ncl = 0;
        for (int s = 0; s < nsub; s++) {
            ESP_CHECK(sub[s].cluster >= 0, ESPRESO_ERR_STATE,
                      "HYBRID_FETI: subdomain " + std::to_string(s) + " has no cluster (espreso_feti_set_cluster)");
            ncl = std::max(ncl, sub[s].cluster + 1);
        }
        cl_subs.assign(ncl, {});
        for (int s = 0; s < nsub; s++) cl_subs[sub[s].cluster].push_back(s);
        for (int c = 0; c < ncl; c++) ESP_CHECK(!cl_subs[c].empty(), ESPRESO_ERR_ARG, "empty cluster id");
    }

    // ---- device: problem arrays ----
    note("load: host preprocessing", t0);
    csr_up.t.join();
    if (csr_err) std::rethrow_exception(csr_err);
    if (measure) {
        ESP_HIP(hipStreamSynchronize(up.s));
        std::fprintf(stderr,
                     "espreso: load: K CSR staging buffers at %.4f s, row pointers queued at %.4f s, last chunk queued "
                     "at %.4f s (host copies %.4f s, waits for the copy engine %.4f s)\n",
                     csr_t[0], csr_t[2], csr_t[1], stage_copy_s, stage_wait_s);
    }
    note("load: K CSR upload joined", t0);
    D.nsub = nsub;
    D.nlam = nlam;
    D.ncoarse = ncoarse;
    D.max_m = max_m;
    D.nbnd = nb;
    std::vector<int> mv(nsub), kd(nsub);
    for (int s = 0; s < nsub; s++) { mv[s] = sub[s].m; kd[s] = sub[s].kdim; }
    D.m = upload(mv);
    D.kdim = upload(kd);
    D.asub = upload(asubv);
    D.boff = upload(boff);
    D.rboff = upload(rboff);
    D.bl_ptr = upload(bl_ptr);
    D.bl_lam = upload(bl_lam);
    D.bl_sign = upload(bl_sign);
    D.bl_sscale = upload(bl_ss);
    D.le_cnt = upload(le_cnt);
    D.le_b = upload(le_b);
    D.le_sub = upload(le_sub);
    D.le_sign = upload(le_sign);
    D.le_sscale = upload(le_ss);
    D.kbb_ptr = upload(kbb_ptr);
    D.kbb_col = upload(kbb_col);
    D.kbb_val = alloc<double>(kbb_pos.size());
    d_kbb_pos = upload(kbb_pos);
    n_kbb = (int64_t)kbb_pos.size();
    if (superdir) {
        d_sib_ptr = upload(sib_ptr);
        d_sib_col = upload(sib_col);
        d_sib_pos = upload(sib_pos);
        d_sdiag_pos = upload(sdiag_pos);
        d_sbi_ptr = upload(sbi_ptr);
        d_sbi_col = upload(sbi_col);
        d_sbi_pos = upload(sbi_pos);
        d_sib_val = alloc<double>(n_sib);
        d_sbi_val = alloc<double>(n_sbi);
        d_sdiag = alloc<double>(n_near);
        d_snear = alloc<double>(n_near);
    }
    D.Rb = upload(Rb);
    D.yb = alloc<double>(nb);
    D.xb = alloc<double>(nb);

    d_fb = upload(fb);
    d_hf = alloc<double>(nb);
    d_ubf = alloc<double>(nb);
    d_e0f = upload(e);
    n_e0s = std::max<int64_t>(1, coffv[nsub]);
    d_e0s = alloc<double>(n_e0s);
    d_alpha_tot = alloc<double>(coffv[nsub]);
    double *d_Q = upload(Q);
    // keep what the coarse layout and the cluster setup need
    Rb_h = std::move(Rb);
    e0f_h = std::move(e);
    le_cnt_h = std::move(le_cnt);
    le_sub_h = std::move(le_sub);
    le_b_h = std::move(le_b);
    le_sign_h = std::move(le_sign);

    // packed pools: one strip geometry per storage precision in use (F_i and
    // S_D may be stored in different precisions)
    ESP_CHECK(opt.sd_precision >= 0 && opt.sd_precision <= 2, ESPRESO_ERR_ARG, "sd_precision must be 0, 1 or 2");
    D.sd_prec = opt.sd_precision;
    std::vector<int64_t> poff[3];
    for (int p = 0; p < 3; p++) {
        if (!prec_used(p)) continue;
        poff[p].assign(nsub + 1, 0);
        for (int s = 0; s < nsub; s++) poff[p][s + 1] = poff[p][s] + strip_size(sub[s].m, strip_rows(p));
        D.poff[p] = upload(poff[p]);
    }
    const std::vector<int64_t> &foff = poff[0];
    const double esize[3] = {8, 4, 2};
    D.Fd = alloc<double>(foff[nsub]);
    D.foff = D.poff[0];
    D.sdoff = D.foff;
    fsc_bytes = (double)poff[prec][nsub] * esize[prec];
    sd_bytes = (double)poff[D.sd_prec][nsub] * esize[D.sd_prec];
    if (prec == 1) D.Ff = alloc<float>(poff[1][nsub]);
    if (prec == 2) D.Fh = alloc<uint16_t>(poff[2][nsub]);
    if (D.sd_prec == 0) D.SD = alloc<double>(foff[nsub]);
    if (D.sd_prec == 1) D.SDf = alloc<float>(poff[1][nsub]);
    if (D.sd_prec == 2) D.SDh = alloc<uint16_t>(poff[2][nsub]);
    D.fscale = alloc<float>(nsub);
    D.sdscale = alloc<float>(nsub);
    D.fnrm = alloc<double>(nsub);
    D.sdnrm = alloc<double>(nsub);
    {
        std::vector<float> ones(nsub, 1.0f);
        ESP_HIP(hipMemcpyAsync(D.fscale, ones.data(), nsub * sizeof(float), hipMemcpyHostToDevice, st));
        ESP_HIP(hipMemcpyAsync(D.sdscale, ones.data(), nsub * sizeof(float), hipMemcpyHostToDevice, st));
    }

    // primal support
    std::vector<int64_t> n_off(nsub + 1, 0), band_off(nsub + 1, 0), yf_off(nsub + 1, 0), r_off(nsub + 1, 0);
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        n_off[s + 1] = n_off[s] + h.n;
        band_off[s + 1] = band_off[s] + (int64_t)h.ni * (h.bw + 1);
        yf_off[s + 1] = yf_off[s] + h.ni;
        r_off[s + 1] = r_off[s] + (int64_t)h.n * h.kdim;
    }
    {
        std::vector<int> niv(nsub), nv(nsub), bwv(nsub);
        for (int s = 0; s < nsub; s++) {
            niv[s] = sub[s].ni;
            nv[s] = sub[s].n;
            bwv[s] = sub[s].bw;
        }
        note("load: small uploads", t0);
        P.nz_off = upload(nz_off);
        P.rp_off = upload(rp_off);
        // the first readers of the K_i CSR arrays follow
        ESP_HIP(hipStreamWaitEvent(st, up.done, 0));
        // structured grid: K_i values from the unit element matrix and the element materials
        if (grid_dev) grid_stiffness();
        gather_prec_values();
        P.perm = upload_pieces<int>(n_off, [&](int s) { return sub[s].perm.data(); });
        P.iperm = upload_pieces<int>(n_off, [&](int s) { return sub[s].iperm.data(); });
        P.R = upload_pieces<double>(r_off, [&](int s) { return sub[s].R.data(); });
        P.ni = upload(niv);
        P.max_ni = nsub ? *std::max_element(niv.begin(), niv.end()) : 0;
        P.n = upload(nv);
        P.bw = upload(bwv);
        P.m = D.m;
        P.kdim = D.kdim;
        P.asub = D.asub;
        P.boff = D.boff;
        P.perm_off = upload(n_off);
        P.band_off = upload(band_off);
        P.yf_off = upload(yf_off);
        P.work_off = P.yf_off;
        P.r_off = upload(r_off);
        P.band = alloc<double>(band_off[nsub]);
        P.yf = alloc<double>(yf_off[nsub]);
        P.work = alloc<double>(yf_off[nsub]);
        nyf = yf_off[nsub];
        P.alpha = d_alpha_tot;
        d_u = alloc<double>(n_off[nsub]);
        uoff = n_off;
        d_f = upload_pieces<double>(n_off, [&](int s) { return sub[s].f.data(); });
        sync();
        staging_release();
    }
    d_Qm = d_Q;
    for (int p = 0; p < 3; p++) h_poff[p] = poff[p];
    h_boff = boff;
    note("load: upload", t0);
    if (hybrid) load_clusters();
    else coarse_layout_subdomains();
    note("load: coarse space, GG^T", t0);

    // ---- SYMV work lists (the item size is tuned at the first set-up) ----
    {
        max_m_h = max_m;
        int64_t target = SYMV_DEFAULT_TARGET;
        if (const char *e = std::getenv("ESPRESO_SYMV_ITEM")) {
            target = std::max<int64_t>(4096, std::atoll(e));
            symv_tuned = true;
        }
        // partial vectors sized for the smallest item size any candidate uses
        std::vector<SymvItem> it;
        std::vector<int> i0, ni;
        int64_t plen_max = 1, pl = 0;
        for (int p = 0; p < 3; p++) {
            if (!prec_used(p)) continue;
            build_items(p, std::min(target, SYMV_TUNE_TARGETS[0]), it, i0, ni, pl);
            plen_max = std::max(plen_max, pl);
        }
        D.ypart = alloc<double>(plen_max);
        for (int p = 0; p < 3; p++)
            if (prec_used(p)) upload_items(p, target);
    }
    // PCG vectors
    d_x = alloc<double>(nlam); d_r = alloc<double>(nlam); d_w = alloc<double>(nlam); d_y = alloc<double>(nlam);
    d_z = alloc<double>(nlam); d_p = alloc<double>(nlam); d_Ap = alloc<double>(nlam); d_b = alloc<double>(nlam);
    d_tmp = alloc<double>(nlam);
    d_r2 = alloc<double>(nlam);
    d_partial = alloc<double>(1024);
    d_scal = alloc<double>(16);
    d_c = upload(c);
    for (int s = 0; s < nsub; s++) kkt_max_n = std::max(kkt_max_n, sub[s].n);
    d_kkt = alloc<double>(2 * nb + nlam + 2 * (int64_t)nsub + 8 + 2 * (int64_t)nsub * kkt_groups(kkt_max_n));
    {
        // algorithmic bytes of one SYMV: the packed triangle in the storage
        // precision, 20 B per B1 entry (index, coefficient, lambda value), y_b
        const double esz[3] = {8, 4, 2};
        double tri = 0, gat = 0, yb = 0;
        for (int s = 0; s < nsub; s++) {
            const double m = sub[s].m;
            tri += m * (m + 1) / 2;
            gat += 20.0 * sub[s].lam.size();
            yb += 8 * m;
        }
        last.symv_bytes = esz[prec] * tri + gat + yb;
        last.symv_bytes_sd = esz[D.sd_prec] * tri + gat + yb;
    }
    check_launches("load");
    sync();
    loaded = true;
    last.time_load = wall() - t0;
    // after st has waited for the upload stream: the scratch of the upload
    // (block-cache images, whose release waits for st) goes back
    ESP_CHECK(!sym_err, ESPRESO_ERR_ARG, "K_i is not structurally symmetric (an entry's mirror is missing)");
    if (d_half) {
        int e = 0;
        ESP_HIP(hipMemcpyAsync(&e, d_half + nsub, sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        ESP_CHECK(!e, ESPRESO_ERR_ARG, "K_i is not structurally symmetric (an entry's mirror is missing)");
        release(d_half);
        release(d_low);
        release(d_rpo);
    }
    if (colidx16) release(colidx16);
    if (d_pos) release(d_pos);
    if (pin_img) release(pin_img);
    if (vpin_img) release(vpin_img);
    note("load: total", t0);
}

// K values on the device for the structured grid: every CSR entry sums the
// element contributions s_e * Ke0 (grid_escale, the unit element matrix)
void espreso_feti_s::grid_stiffness() {
    int64_t max_nnz = 0;
    for (int s = 0; s < nsub; s++) max_nnz = std::max<int64_t>(max_nnz, h_nzoff[s + 1] - h_nzoff[s]);
    int *d_box = upload(grid_box);
    double *d_Ke = upload(gridKe0), *d_esc = upload(grid_escale);
    launch_grid_stiffness(nsub, ndof, d_box, gridNE, d_Ke, d_esc, P.rowptr, P.rp_off, P.nz_off, (int)max_nnz, P.vals,
                          st);
    sync();
    for (void *q : {(void *)d_box, (void *)d_Ke, (void *)d_esc}) release(q);
}

// preconditioner values taken from K: K_bb (lumped) and, for SUPER_DIRICHLET,
// K_ib, K_bi and diag(K_ii) of the near-interior DOFs
void espreso_feti_s::gather_prec_values() {
    launch_gather_values(n_kbb, d_kbb_pos, P.vals, D.kbb_val, st);
    if (opt.preconditioner == ESPRESO_PREC_SUPER_DIRICHLET) {
        launch_gather_values(n_sib, d_sib_pos, P.vals, d_sib_val, st);
        launch_gather_values(n_sbi, d_sbi_pos, P.vals, d_sbi_val, st);
        launch_gather_values(n_near, d_sdiag_pos, P.vals, d_sdiag, st);
    }
    values_dirty = false;
}

// GPU part of the setup, repeatable on the resident inputs: local dual
// operators and Dirichlet blocks, coarse inverse, right-hand side b.
void espreso_feti_s::gpu_setup() {
    ESP_CHECK(loaded, ESPRESO_ERR_STATE, "inputs not loaded");
    drop_graphs();
    if (const char *e = std::getenv("ESPRESO_HIP_GRAPHS")) use_graphs = std::atoi(e) != 0;
    ready = false;
    double ta = wall();
    if (values_dirty) gather_prec_values();
    assemble(d_f, d_Qm, h_poff, h_boff);
    check_launches("assembly");
    last.time_assembly = wall() - ta;
    note("setup: assembly", ta);
    double tc = wall();
    if (hybrid) cluster_setup();
    join_ggt();
    coarse();
    compute_b();
    check_launches("coarse set-up");
    sync();
    last.time_coarse = wall() - tc;
    note("setup: cluster, coarse, b", tc);
    tune_symv();  // first set-up only (outside the timed phases)
    check_launches("autotuner");
    ready = true;
}

// K^+ f on the boundary: ubf = F_i hf (FP64); hybrid: + cluster correction;
// then b = B K^+ f - c (cluster.cpp:2660-2769)
void espreso_feti_s::compute_b() { compute_b_into(d_hf, d_ubf, d_e0f, d_b); }

void espreso_feti_s::compute_b_into(const double *hf, double *ubf, const double *e0f, double *b) {
    launch_symv(D, false, 0, hf, 0, false, st);
    ESP_HIP(hipMemcpyAsync(ubf, D.yb, D.nbnd * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (hybrid) launch_cluster_correction(D, CL, e0f, D.yb, 1, st);
    launch_lambda_gather(D, D.yb, false, b, st);
    launch_axpby(b, b, d_c, nlam, 1.0, -1.0, st);
}

// host part of new loads f (all subdomains, local order): e0f = -R_s^T f_s per
// subdomain and the coarse right-hand side e (hybrid: per cluster through N_s)
void espreso_feti_s::rhs_host(const double *f, hvec<double> &e0f, std::vector<double> &e) const {
    e0f.assign(asubv[nsub], 0.0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        const double *fs = f + uoff[s];
        for (int a = 0; a < h.kdim; a++) {
            double acc = 0;
            for (int i = 0; i < h.n; i++) acc += h.R[i + (size_t)a * h.n] * fs[i];
            e0f[asubv[s] + a] = -acc;
        }
    }
    if (hybrid) {
        e.assign(ncoarse, 0.0);
        for (int s = 0; s < nsub; s++) {
            const int c = cblk_h[s], kn = (int)((noff_h[s + 1] - noff_h[s]) / std::max(1, sub[s].kdim));
            for (int b = 0; b < kn; b++)
                for (int a = 0; a < sub[s].kdim; a++)
                    e[ccoff_h[c] + b] += Ns_h[noff_h[s] + a + (int64_t)b * sub[s].kdim] * e0f[asubv[s] + a];
        }
    } else {
        e.assign(e0f.begin(), e0f.end());
    }
}

// New loads (host sub[s].f) on the assembled operators: e = -R^T f (hybrid:
// per cluster through N_s), y_f and hf by interior band solves, b.
void espreso_feti_s::resolve_rhs() {
    std::vector<double> ff(uoff[nsub]);
    for (int s = 0; s < nsub; s++) std::copy(sub[s].f.begin(), sub[s].f.end(), ff.begin() + uoff[s]);
    std::vector<double> e;
    rhs_host(ff.data(), e0f_h, e);
    ESP_HIP(hipMemcpyAsync(d_f, ff.data(), ff.size() * sizeof(double), hipMemcpyHostToDevice, st));
    ESP_HIP(hipMemcpyAsync(d_e0f, e0f_h.data(), e0f_h.size() * sizeof(double), hipMemcpyHostToDevice, st));
    ESP_HIP(hipMemcpyAsync(d_e, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, st));
    launch_rhs_prepare(P, nsub, d_f, d_hf, st);
    compute_b();
    sync();
}

void espreso_feti_s::assemble(const double *d_f, const double *d_Q, const std::vector<int64_t> *poff,
                              const std::vector<int64_t> &boff) {
    const int prec = opt.sc_precision;
    // per-subdomain workspace (doubles): K_ii in band storage, [K_ib | f], S,
    // W (in [K_ib | f]'s storage when it fits there), diagonal inverses
    auto kii_rows = [](const SubHost &h) { return std::min<int64_t>((int64_t)(h.band_tiles + 2) * TILE, h.ni + TILE); };
    auto w_apart = [](const SubHost &h) { return (int64_t)h.m * h.m > (int64_t)h.ni * (h.m + 1); };
    auto ws = [&](const SubHost &h) {
        const int64_t nt = std::max((h.ni + TILE - 1) / TILE, (h.m + TILE - 1) / TILE);
        return (int64_t)h.ni * kii_rows(h) + (int64_t)h.ni * (h.m + 1) + (int64_t)h.m * h.m +
               (w_apart(h) ? (int64_t)h.m * h.m : 0) + nt * TILE * TILE;
    };
    // wave plan and workspace: planned once at the first set-up and kept
    // (re-planning and reallocating a multi-GB workspace costs more than the
    // assembly).  The momentary free memory can lag far behind what an earlier
    // instance gave back (the driver reclaims large frees lazily, and the pool
    // keeps what it holds), so the plan also counts on the device memory this
    // instance does not hold; when that is not there the workspace allocation
    // fails and the plan is redone with half the budget.
    if (wave_plan.empty()) {
        size_t freeb = 0, totalb = 0;
        ESP_HIP(hipMemGetInfo(&freeb, &totalb));
        double pool_free = (double)workspace_cache().idle(opt.device);
        {
            hipMemPool_t pool = nullptr;
            uint64_t reserved = 0, used = 0;
            ESP_HIP(hipDeviceGetDefaultMemPool(&pool, opt.device));
            ESP_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved));
            ESP_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used));
            if (reserved > used) pool_free += (double)(reserved - used);
        }
        double budget = std::max(0.85 * ((double)freeb + pool_free), 0.85 * (double)totalb - (double)held_bytes);
        if (opt.memory_budget_gb > 0) budget = std::min(budget, opt.memory_budget_gb * 1e9);
        double need = 0, largest = 0;
        for (int s = 0; s < nsub; s++) {
            need += 8.0 * ws(sub[s]);
            largest = std::max(largest, 8.0 * ws(sub[s]));
        }
        budget = std::min(budget, need);
        for (;;) {
            ESP_CHECK(largest <= budget, ESPRESO_ERR_MEMORY,
                      "one subdomain's assembly workspace exceeds the device memory left");
            wave_plan.clear();
            size_t most = 0;
            for (int a = 0; a < nsub;) {
                int b = a;
                double bytes = 0;
                while (b < nsub && (b == a || bytes + 8.0 * ws(sub[b]) <= budget)) bytes += 8.0 * ws(sub[b++]);
                wave_plan.push_back(b);
                most = std::max(most, (size_t)(bytes / 8) + 1);
                a = b;
            }
            ESP_HIP(hipStreamSynchronize(st));  // a lent block may have been the last instance's
            asm_pool = (double *)workspace_cache().acquire(opt.device, most * sizeof(double));
            if (asm_pool) {
                asm_pool_n = most;
                break;
            }
            budget *= 0.5;
        }
        if (measure)
            std::fprintf(stderr, "espreso: assembly workspace %.1f GB in %zu wave(s) (free %.1f GB, pool %.1f GB, "
                                 "held %.1f GB)\n", 8e-9 * asm_pool_n, wave_plan.size(), 1e-9 * freeb,
                         1e-9 * pool_free, 1e-9 * held_bytes);
    }
    int s0 = 0;
    asm_flops = 0;
    for (const int s1 : wave_plan) {
        const int nbt = s1 - s0;
        double *pool = asm_pool;
        std::vector<DenseDesc> desc(nbt);
        std::vector<int> cs_all;
        std::vector<int64_t> cs_off(nbt);
        int max_ntii = 0, max_ntm = 0, max_ntc = 0, max_band = 0, max_n = 0, max_m = 0, max_ni = 0;
        double *cur = pool;
        for (int b = 0; b < nbt; b++) {
            const SubHost &h = sub[s0 + b];
            DenseDesc &d = desc[b];
            d.ni = h.ni;
            d.m = h.m;
            d.ntii = (h.ni + TILE - 1) / TILE;
            d.ntm = (h.m + TILE - 1) / TILE;
            d.ntc = (h.m + 1 + TILE - 1) / TILE;
            d.band_ii = h.band_tiles;
            d.ldk = (int)kii_rows(h) - 1;
            d.Kii = cur + (TILE - 1); cur += (int64_t)h.ni * kii_rows(h);
            d.Kib = cur; cur += (int64_t)h.ni * (h.m + 1);
            d.S = cur; cur += (int64_t)h.m * h.m;
            if (w_apart(h)) { d.W = cur; cur += (int64_t)h.m * h.m; }
            else d.W = d.Kib;
            d.Dinv = cur; cur += (int64_t)std::max(d.ntii, d.ntm) * TILE * TILE;
            cs_off[b] = (int64_t)cs_all.size();
            cs_all.insert(cs_all.end(), h.colstart.begin(), h.colstart.end());
            max_ntii = std::max(max_ntii, d.ntii);
            max_ntm = std::max(max_ntm, d.ntm);
            max_ntc = std::max(max_ntc, d.ntc);
            max_band = std::max(max_band, d.band_ii);
            max_n = std::max(max_n, h.n);
            max_m = std::max(max_m, h.m);
            max_ni = std::max(max_ni, h.ni);
            asm_flops += tile_flops(h);
        }
        int *d_cs = upload(cs_all);
        int *d_info = alloc<int>(nbt);
        ESP_HIP(hipMemsetAsync(d_info, 0, nbt * sizeof(int), st));
        for (int b = 0; b < nbt; b++) {
            desc[b].colstart = d_cs + cs_off[b];
            desc[b].info = d_info + b;
        }
        DenseDesc *d_desc = upload(desc);
        // wave-local offset tables
        std::vector<int64_t> rp_off(nbt), nz_off(nbt), n_off(nbt), b_off(nbt), q_off(nbt), p_off[3], band_off(nbt),
            yf_off(nbt);
        for (int p = 0; p < 3; p++)
            if (!poff[p].empty()) p_off[p].assign(poff[p].begin() + s0, poff[p].begin() + s1);
        std::vector<int> nloc(nbt), kd(nbt), bw(nbt);
        {
            // global offsets of this wave's subdomains in the persistent pools
            int64_t rp = 0, nz = 0, nn = 0, bo = 0, qo = 0;
            for (int s = 0; s < s0; s++) {
                rp += sub[s].n + 1;
                nz += sub[s].rowptr[sub[s].n];
                nn += sub[s].n;
                qo += (int64_t)sub[s].m * sub[s].kdim;
            }
            int64_t band = 0, yf = 0;
            for (int s = 0; s < s0; s++) { band += (int64_t)sub[s].ni * (sub[s].bw + 1); yf += sub[s].ni; }
            for (int b = 0; b < nbt; b++) {
                const SubHost &h = sub[s0 + b];
                rp_off[b] = rp; nz_off[b] = nz; n_off[b] = nn; b_off[b] = boff[s0 + b]; q_off[b] = qo;
                band_off[b] = band; yf_off[b] = yf;
                rp += h.n + 1; nz += h.rowptr[h.n]; nn += h.n; qo += (int64_t)h.m * h.kdim;
                band += (int64_t)h.ni * (h.bw + 1); yf += h.ni;
                nloc[b] = h.n; kd[b] = h.kdim; bw[b] = h.bw;
                (void)bo;
            }
        }
        int64_t *d_rp_off = upload(rp_off), *d_nz_off = upload(nz_off), *d_n_off = upload(n_off),
                *d_b_off = upload(b_off), *d_q_off = upload(q_off), *d_band_off = upload(band_off),
                *d_yf_off = upload(yf_off);
        int64_t *d_p_off[3] = {nullptr, nullptr, nullptr};
        for (int p = 0; p < 3; p++)
            if (!p_off[p].empty()) d_p_off[p] = upload(p_off[p]);
        int *d_nloc = upload(nloc), *d_kd = upload(kd), *d_bw = upload(bw);
        double *d_rho = alloc<double>(nbt);

        const double tw = wall();
        if (measure) sync();
        note("assembly: wave inputs", tw);
        launch_materialize(d_desc, nbt, P.rowptr, d_rp_off, P.colidx, P.vals, d_nz_off, P.iperm, P.perm, d_n_off, d_f,
                           d_nloc, max_n, st);
        if (measure) sync();
        note("assembly: materialize", tw);
        factor_solve(d_desc, nbt, 0, max_ntii, max_band, max_ntc);
        if (measure) sync();
        note("assembly: K_ii factor, solve", tw);
        // h_f and the band copy of the K_ii factor read only [K_ib f] and K_ii:
        // on asm_st beside S = K_bb - Y^T Y (the S-stage substitution that
        // later writes W over [K_ib f] runs behind them on asm_st)
        asm_fork();
        launch_hf(d_desc, nbt, d_fb, d_hf, d_b_off, max_m, asm_st);
        launch_band_extract(d_desc, nbt, P.band, d_band_off, d_bw, P.yf, d_yf_off, max_ni, asm_st);
        launch_syrk(d_desc, nbt, 0, max_ntm, st);
        launch_add_kbb(d_desc, nbt, P.rowptr, d_rp_off, P.colidx, P.vals, d_nz_off, P.iperm, P.perm, d_n_off, max_m,
                       st);
        // S_D = S = K_bb - Y^T Y: its row norms (asm_st) and its packed copy (st) both only
        // read S; the kernel regularisation then changes S
        asm_fork();
        launch_rownorm(d_desc, nbt, D.sdnrm + s0, max_m, asm_st);
        if (D.sd_prec == 0) launch_pack(d_desc, nbt, 0, D.SD, d_p_off[0], nullptr, max_m, st);
        if (D.sd_prec == 1) launch_pack(d_desc, nbt, 1, D.SDf, d_p_off[1], nullptr, max_m, st);
        if (D.sd_prec == 2) launch_pack(d_desc, nbt, 2, D.SDh, d_p_off[2], D.sdscale + s0, max_m, st);
        asm_join();
        launch_kernel_reg(d_desc, nbt, d_Q, d_q_off, d_kd, d_rho, max_m, st);
        if (hybrid)
            ESP_HIP(hipMemcpyAsync(CL.rho_s + s0, d_rho, nbt * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (measure) sync();
        note("assembly: S_D, band", tw);
        factor_solve(d_desc, nbt, 1, max_ntm, max_ntm, max_ntm);
        if (measure) sync();
        note("assembly: S factor, inverse", tw);
        launch_syrk(d_desc, nbt, 1, max_ntm, st);
        if (measure) sync();
        note("assembly: F = W^T W", tw);
        // F's row norms beside its packed copies (both only read F); joined below
        asm_fork();
        launch_rownorm(d_desc, nbt, D.fnrm + s0, max_m, asm_st);
        if (hybrid) {
            std::vector<int> nq(nq_h.begin() + s0, nq_h.begin() + s1);
            std::vector<int64_t> qo(nbt), fo(nbt), co(nbt);
            int max_nq = 0;
            int64_t fb = 0;
            for (int s = 0; s < s0; s++) fb += (int64_t)sub[s].m * nq_h[s];
            for (int b = 0; b < nbt; b++) {
                qo[b] = qoff_h[s0 + b];
                co[b] = fccoff_h[s0 + b];
                fo[b] = fb;
                fb += (int64_t)sub[s0 + b].m * nq_h[s0 + b];
                max_nq = std::max(max_nq, nq[b]);
            }
            int *d_nq = upload(nq);
            int64_t *d_qo = upload(qo), *d_fo = upload(fo), *d_co = upload(co);
            if (max_nq > 0)
                launch_extract_b0(d_desc, nbt, d_nq, d_qo, CL.qe_ptr, CL.qe_j, CL.qe_c, d_fo, d_co, CL.FB0, d_Fcc, max_m,
                                  max_nq, max_qe_h > 1, st);
            sync();
            for (void *p : {(void *)d_nq, (void *)d_qo, (void *)d_fo, (void *)d_co}) release(p);
        }
        launch_pack(d_desc, nbt, 0, D.Fd, d_p_off[0], nullptr, max_m, st);
        if (prec == 1) launch_pack(d_desc, nbt, 1, D.Ff, d_p_off[1], nullptr, max_m, st);
        if (prec == 2) launch_pack(d_desc, nbt, 2, D.Fh, d_p_off[2], D.fscale + s0, max_m, st);
        asm_join();
        if (measure) sync();
        note("assembly: pack", tw);
        std::vector<int> info(nbt);
        ESP_HIP(hipMemcpyAsync(info.data(), d_info, nbt * sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        note("assembly: status", tw);
        for (int b = 0; b < nbt; b++)
            ESP_CHECK(info[b] == 0, ESPRESO_ERR_NUMERIC,
                      "dense factorisation breakdown in subdomain " + std::to_string(s0 + b) + " at column " +
                          std::to_string(info[b]));
        for (void *p : {(void *)d_cs, (void *)d_info, (void *)d_desc, (void *)d_rp_off, (void *)d_nz_off,
                        (void *)d_n_off, (void *)d_b_off, (void *)d_q_off, (void *)d_band_off, (void *)d_yf_off,
                        (void *)d_nloc, (void *)d_kd, (void *)d_bw, (void *)d_rho})
            release(p);
        for (int64_t *q : d_p_off)
            if (q) release(q);
        note("assembly: wave released", tw);
        s0 = s1;
    }
    if (hybrid && CL.FB0f) launch_fb0_round(D, CL, st);
}

void espreso_feti_s::coarse() {
    const int nc = ncoarse;
    const int nt = (nc + TILE - 1) / TILE;
    if (D.GGinv) release(D.GGinv);
    D.GGinv = nullptr;
    double *S = alloc<double>((size_t)nc * nc);
    ESP_HIP(hipMemcpyAsync(S, d_GGt, (size_t)nc * nc * sizeof(double), hipMemcpyDeviceToDevice, st));
    double *W = alloc<double>((size_t)nc * nc);
    double *Dinv = alloc<double>((size_t)nt * TILE * TILE);
    int *info = alloc<int>(1);
    ESP_HIP(hipMemsetAsync(info, 0, sizeof(int), st));
    DenseDesc d{};
    d.S = S;
    d.W = W;
    d.Dinv = Dinv;
    d.info = info;
    d.ni = 0;
    d.m = nc;
    d.ntm = nt;
    d.ntc = nt;
    DenseDesc *dd = upload(std::vector<DenseDesc>{d});
    // one matrix: right-looking Cholesky steps (the trailing update over all
    // tiles) and W = L^-1 by recursive doubling (two launches per level over
    // all tiles) instead of left-looking panels and one substitution per row
    launch_potrf_rl(dd, 1, nt, st);
    launch_trtri(dd, 1, nt, st);
    launch_syrk(dd, 1, 1, nt, st);
    launch_symmetrize(S, nc, st);
    int hinfo = 0;
    ESP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    ESP_CHECK(hinfo == 0, ESPRESO_ERR_NUMERIC, "GG^T factorisation breakdown");
    D.GGinv = S;
    release(W);
    release(Dinv);
    release(info);
    release(dd);
}

// Left-looking block Cholesky of A (which 0: K_ii, 1: S) on the main stream and
// the forward substitution X = L^-1 B on asm_st.  Block row k of the substitution
// needs the diagonal inverse of step k and the factor's rows k of earlier steps,
// all done once the diagonal kernel of step k is; it writes only B, which the
// Cholesky steps do not touch.  So it starts behind an event after that kernel
// and fills the GPU while the later, latency-bound diagonal steps run.
void espreso_feti_s::factor_solve(const DenseDesc *desc, int nb, int which, int nt, int band, int ntc) {
    if (!asm_st) ESP_HIP(hipStreamCreate(&asm_st));
    while ((int)asm_ev.size() < nt + 1) {
        hipEvent_t e = nullptr;
        ESP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        asm_ev.push_back(e);
    }
    // the second stream starts behind everything the main stream has queued
    ESP_HIP(hipEventRecord(asm_ev[nt], st));
    ESP_HIP(hipStreamWaitEvent(asm_st, asm_ev[nt], 0));
    for (int k = 0; k < nt; k++) {
        launch_potrf_diag(desc, nb, which, k, st);
        ESP_HIP(hipEventRecord(asm_ev[k], st));
        launch_potrf_panel(desc, nb, which, k, band, st);
        ESP_HIP(hipStreamWaitEvent(asm_st, asm_ev[k], 0));
        launch_trsm_step(desc, nb, which, k, band, ntc, asm_st);
    }
    ESP_HIP(hipEventRecord(asm_ev[nt], asm_st));
    ESP_HIP(hipStreamWaitEvent(st, asm_ev[nt], 0));
}

// Flops of the assembly tiles of one subdomain, following the kernels' own tile
// rules (band windows, the leading zero tiles of [K_ib f] that the TRSM and SYRK
// skip, triangular ranges in the S chain), partial tiles at their real size:
// the figure the reported assembly rate is computed from.  MFMA instructions
// issue slightly more (partial tiles are padded to 16).
double espreso_feti_s::tile_flops(const SubHost &h) const {
    double f = 0;
    for (int which = 0; which < 2; which++) {
        const int n = which ? h.m : h.ni, nt = (n + TILE - 1) / TILE;
        const int band = which ? nt : h.band_tiles;
        const int ncols = which ? h.m : h.m + 1, ntc = (ncols + TILE - 1) / TILE;
        for (int k = 0; k < nt; k++) {
            const double k0 = (double)k * TILE, nb = std::min(TILE, n - k * TILE);
            const double j0 = (double)std::max(0, k - band) * TILE;
            f += 2 * nb * nb * (k0 - j0) + 2.0 / 3.0 * nb * nb * nb;  // diagonal update, factor and inverse
            for (int i = k + 1; i < nt && i <= k + band; i++) {        // panel
                const double mi = std::min(TILE, n - i * TILE), w0 = (double)std::max(0, i - band) * TILE;
                f += 2 * mi * nb * std::max(0.0, k0 - w0) + 2 * mi * nb * nb;
            }
            for (int J = 0; J < ntc; J++) {  // forward substitution, block row k
                const int cs = which ? J : h.colstart[J];
                if (cs > k || (which && J == k)) continue;
                const double nj = std::min(TILE, ncols - J * TILE), w0 = (double)std::max(cs, k - band) * TILE;
                f += 2 * nb * nj * std::max(0.0, k0 - w0) + 2 * nb * nj * nb;
            }
        }
        const int ntm = (h.m + TILE - 1) / TILE;
        for (int I = 0; I < ntm; I++)  // S -= Y^T Y  or  F = W^T W, lower tiles
            for (int J = 0; J <= I; J++) {
                const double mi = std::min(TILE, h.m - I * TILE), mj = std::min(TILE, h.m - J * TILE);
                const double ks = which ? (double)I * TILE : (double)std::max(h.colstart[I], h.colstart[J]) * TILE;
                const double len = (which ? h.m : h.ni) - ks;
                if (len > 0) f += 2 * mi * mj * len;
            }
    }
    return f;
}

void espreso_feti_s::build_items(int prec, int64_t target, std::vector<SymvItem> &items, std::vector<int> &item0,
                                 std::vector<int> &nit, int64_t &plen) {
    const int R = strip_rows(prec);
    items.clear();
    item0.assign(nsub, 0);
    nit.assign(nsub, 0);
    plen = 0;
    for (int s = 0; s < nsub; s++) {
        const int m = sub[s].m;
        const int nst = (int)strip_count(m, R);
        const int64_t tot = strip_size(m, R);
        int want = (int)std::max<int64_t>(1, std::min<int64_t>(nst, (tot + target - 1) / target));
        item0[s] = (int)items.size();
        int t = 0;
        for (int k = 0; k < want && t < nst; k++) {
            const int64_t goal = tot * (k + 1) / want;
            int t1 = t + 1;
            while (t1 < nst && strip_offset(m, R, t1) < goal) t1++;
            if (k == want - 1) t1 = nst;
            SymvItem it{};
            it.sub = s;
            it.t0 = t;
            it.t1 = t1;
            it.part_off = plen;
            plen += m - t * R;
            items.push_back(it);
            t = t1;
        }
        nit[s] = (int)items.size() - item0[s];
    }
}

// (re)build and upload the SYMV work list of a storage precision
void espreso_feti_s::upload_items(int prec, int64_t target) {
    std::vector<SymvItem> it;
    std::vector<int> i0, ni;
    int64_t plen = 0;
    build_items(prec, target, it, i0, ni, plen);
    for (void *q : {(void *)D.items[prec], (void *)D.item0[prec], (void *)D.nitems[prec]})
        if (q) release(q);
    D.items[prec] = upload(it);
    D.item0[prec] = upload(i0);
    D.nitems[prec] = upload(ni);
    D.n_items[prec] = (int)it.size();
    symv_target[prec] = target;
}

// Autotuner of the SYMV work-item size, keyed on the subdomain size: at the
// first set-up every candidate is timed on the assembled F_i (HIP events, 2
// applies each after one warm-up apply) and the fastest kept; the choice is cached for the process by
// (storage precision, largest interface size rounded to 64).  ESPRESO_SYMV_ITEM
// fixes the size, ESPRESO_AUTOTUNE=0 keeps the default.
void espreso_feti_s::tune_symv() {
    if (symv_tuned) return;
    symv_tuned = true;
    if (const char *e = std::getenv("ESPRESO_AUTOTUNE"))
        if (std::atoi(e) == 0) return;
    static std::mutex mu;
    static std::map<std::pair<int, int>, int64_t> cache;
    for (int prec = 0; prec < 3; prec++) {
        if (!prec_used(prec)) continue;
        const std::pair<int, int> key{prec, (max_m_h + 63) / 64};
        {
            std::lock_guard<std::mutex> lk(mu);
            auto f = cache.find(key);
            if (f != cache.end()) {
                if (f->second != symv_target[prec]) upload_items(prec, f->second);
                continue;
            }
        }
        // time on the F_i of this precision if stored in it, else on S_D
        const bool sd = prec != 0 && prec != opt.sc_precision;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ESP_HIP(hipEventCreate(&e0));
        ESP_HIP(hipEventCreate(&e1));
        int64_t best = symv_target[prec];
        float best_ms = 1e30f;
        launch_symv(D, sd, prec, d_b, 1, false, st);  // warm
        for (const int64_t target : SYMV_TUNE_TARGETS) {
            upload_items(prec, target);
            ESP_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < 2; r++) launch_symv(D, sd, prec, d_b, 1, false, st);
            ESP_HIP(hipEventRecord(e1, st));
            ESP_HIP(hipEventSynchronize(e1));
            float ms = 0;
            ESP_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = target;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        upload_items(prec, best);
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = best;
    }
}

// ------------------------------------------------------ dual operators ----
void espreso_feti_s::apply_F(const double *x, double *y, int prec) {
    launch_symv(D, false, prec, x, 1, false, st);
    if (hybrid) {
        // K~^+ = K^+ - K^+ B0^T lambda0 + R alpha (cluster.cpp:796-916), e0 = -R^T B^T x
        launch_subdomain_kernel_dots(D, x, d_e0s, st);
        launch_cluster_correction(D, CL, d_e0s, D.yb, 1, st, prec != 0);
    }
    launch_lambda_gather(D, D.yb, false, y, st);
}

void espreso_feti_s::apply_prec(const double *x, double *y) {
    switch (opt.preconditioner) {
    case ESPRESO_PREC_NONE:
        ESP_HIP(hipMemcpyAsync(y, x, nlam * sizeof(double), hipMemcpyDeviceToDevice, st));
        break;
    case ESPRESO_PREC_LUMPED:
        launch_boundary_gather(D, x, true, D.xb, st);
        launch_kbb(D, D.xb, D.yb, st);
        launch_lambda_gather(D, D.yb, true, y, st);
        break;
    case ESPRESO_PREC_WEIGHT:
        launch_boundary_gather(D, x, true, D.xb, st);
        launch_lambda_gather(D, D.xb, true, y, st);
        break;
    case ESPRESO_PREC_DIRICHLET:
        launch_symv(D, true, 0, x, 1, true, st);
        launch_lambda_gather(D, D.yb, true, y, st);
        break;
    case ESPRESO_PREC_SUPER_DIRICHLET:
        launch_boundary_gather(D, x, true, D.xb, st);
        launch_superdir(D, n_near, d_sib_ptr, d_sib_col, d_sib_val, d_sdiag, d_sbi_ptr, d_sbi_col, d_sbi_val, d_snear,
                        D.xb, D.yb, st);
        launch_lambda_gather(D, D.yb, true, y, st);
        break;
    default:
        throw Error(ESPRESO_ERR_ARG, "unknown preconditioner");
    }
}

// Projected PCG (itersolver.cpp:1537-1684) on device vectors.  r holds the
// initial residual on entry.  Scalars stay on the device; the host reads
// ||w|| once per iteration for the stop test.
// One iteration of Solve_RegCG (itersolver.cpp:1537-1684) as a launch sequence:
// w = P r, y = P M w, beta, p, Ap = F p, alpha, x, r, ||w||^2.
void espreso_feti_s::pcg_step(double *x, double *r, int prec, bool first) {
    double *s_yw = d_scal + 0, *s_yw_prev = d_scal + 1, *s_pAp = d_scal + 2, *s_ww = d_scal + 3;
    launch_project(D, r, d_w, 0, st);
    if (opt.preconditioner == ESPRESO_PREC_NONE) {
        ESP_HIP(hipMemcpyAsync(d_y, d_w, nlam * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
        apply_prec(d_w, d_z);
        launch_project(D, d_z, d_y, 0, st);
    }
    if (!first) launch_copy_scalar(s_yw, s_yw_prev, st);
    launch_dot(d_y, d_w, nlam, d_partial, s_yw, st);
    launch_p_update(d_p, d_y, nlam, s_yw, s_yw_prev, first, st);
    apply_F(d_p, d_Ap, prec);
    launch_dot(d_p, d_Ap, nlam, d_partial, s_pAp, st);
    launch_xr_update(x, r, d_p, d_Ap, nlam, s_yw, s_pAp, st);
    launch_dot(d_w, d_w, nlam, d_partial, s_ww, st);
}

// iterations after the first replay a HIP graph of pcg_step, captured once
// per (x, r, storage precision); graphs hold device pointers of the current
// set-up and are dropped when the set-up is redone
void espreso_feti_s::pcg_step_graph(double *x, double *r, int prec) {
    for (const StepGraph &g : step_graphs)
        if (g.x == x && g.r == r && g.prec == prec) {
            ESP_HIP(hipGraphLaunch(g.exec, st));
            return;
        }
    hipGraph_t graph = nullptr;
    ESP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        pcg_step(x, r, prec, false);
    } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    ESP_HIP(hipStreamEndCapture(st, &graph));
    StepGraph g{x, r, prec, nullptr};
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    ESP_HIP(e);
    step_graphs.push_back(g);
    ESP_HIP(hipGraphLaunch(g.exec, st));
}

void espreso_feti_s::drop_graphs() {
    for (StepGraph &g : step_graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    step_graphs.clear();
    for (BlockGraph &g : block_graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    block_graphs.clear();
}

void espreso_feti_s::pcg(const double *b, double *x, double *r, int prec, int &iters, double &norm, double &tol,
                         bool use_abs, double tol_abs, double *norm_b) {
    double *s_ww = d_scal + 3, *s_tmp = d_scal + 4;
    // tol = precision * min(||P r0||, ||b||)
    launch_project(D, r, d_w, 0, st);
    launch_dot(d_w, d_w, nlam, d_partial, s_tmp, st);
    launch_dot(b, b, nlam, d_partial, s_tmp + 1, st);
    double hv[2];
    ESP_HIP(hipMemcpyAsync(hv, s_tmp, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    sync();
    tol = use_abs ? tol_abs : opt.precision * std::min(std::sqrt(hv[0]), std::sqrt(hv[1]));
    if (norm_b) *norm_b = std::sqrt(hv[1]);
    iters = 0;
    norm = 0;
    for (int it = 0; it < opt.max_iterations; it++) {
        if (it == 0 || !use_graphs) pcg_step(x, r, prec, it == 0);
        else pcg_step_graph(x, r, prec);
        double ww = 0;
        ESP_HIP(hipMemcpyAsync(&ww, s_ww, sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        ESP_CHECK(std::isfinite(ww), ESPRESO_ERR_NUMERIC, "PCG produced a non-finite residual");
        norm = std::sqrt(ww);
        iters = it + 1;
        if (norm < tol) break;
    }
}

void espreso_feti_s::solve(espreso_feti_stats *out) {
    ESP_CHECK(ready, ESPRESO_ERR_STATE, "espreso_feti_setup not called");
    double t0 = wall();
    // lambda0 = G^T (GG^T)^-1 e ; r = b - F lambda0
    launch_project(D, d_e, d_x, 1, st);
    apply_F(d_x, d_Ap, 0);
    launch_axpby(d_r, d_b, d_Ap, nlam, 1.0, -1.0, st);
    int iters = 0;
    double norm = 0, tol = 0;
    const int prec = opt.sc_precision;
    int refine = 0;
    bool converged = true;
    if (prec == 0) {
        pcg(d_b, d_x, d_r, 0, iters, norm, tol, false, 0, &last.norm_b);
        converged = norm < tol;
    } else {
        // mixed precision: inner PCG on the low-precision F_i, outer FP64
        // residual r = b - F lambda with the FP64 F_i (iterative refinement)
        launch_project(D, d_r, d_w, 0, st);
        launch_dot(d_w, d_w, nlam, d_partial, d_scal + 4, st);
        launch_dot(d_b, d_b, nlam, d_partial, d_scal + 5, st);
        double hv[2];
        ESP_HIP(hipMemcpyAsync(hv, d_scal + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        const double tol64 = opt.precision * std::min(std::sqrt(hv[0]), std::sqrt(hv[1]));
        last.norm_b = std::sqrt(hv[1]);
        double *dx = d_tmp;
        int total = 0;
        for (refine = 0; refine < 20; refine++) {
            // correction: solve P F P dx = P r with the low-precision operator
            ESP_HIP(hipMemsetAsync(dx, 0, nlam * sizeof(double), st));
            double *rr = d_r2;  // inner residual
            ESP_HIP(hipMemcpyAsync(rr, d_r, nlam * sizeof(double), hipMemcpyDeviceToDevice, st));
            int it = 0;
            double nn = 0, tt = 0;
            // inner tolerance: relative 1e-5 of the current projected residual, floor tol64
            launch_project(D, d_r, d_w, 0, st);
            launch_dot(d_w, d_w, nlam, d_partial, d_scal + 6, st);
            double cur = 0;
            ESP_HIP(hipMemcpyAsync(&cur, d_scal + 6, sizeof(double), hipMemcpyDeviceToHost, st));
            sync();
            cur = std::sqrt(cur);
            if (cur < tol64) break;
            pcg(d_r, dx, rr, prec, it, nn, tt, true, std::max(tol64 * 0.5, cur * 1e-5));
            total += it;
            launch_axpby(d_x, d_x, dx, nlam, 1.0, 1.0, st);
            apply_F(d_x, d_Ap, 0);
            launch_axpby(d_r, d_b, d_Ap, nlam, 1.0, -1.0, st);
        }
        launch_project(D, d_r, d_w, 0, st);
        launch_dot(d_w, d_w, nlam, d_partial, d_scal + 6, st);
        double fin = 0;
        ESP_HIP(hipMemcpyAsync(&fin, d_scal + 6, sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        iters = total;
        norm = std::sqrt(fin);
        tol = tol64;
        converged = norm < tol64;
    }
    check_launches("PCG");
    double tp = wall();
    last.time_pcg = tp - t0;
    // amplitudes alpha = (GG^T)^-1 G r ; primal u = K~^+ (f - B^T lambda) + R alpha
    launch_project(D, d_r, d_alpha_c, 2, st);
    // u_b = K^+ f|_b - F_i B^T lambda
    launch_symv(D, false, 0, d_x, 1, false, st);
    launch_axpby(D.xb, d_ubf, D.yb, D.nbnd, 1.0, -1.0, st);
    if (hybrid) {
        // e0 = -R^T (f - B^T lambda) = e0f - (-R^T B^T lambda)
        launch_subdomain_kernel_dots(D, d_x, d_e0s, st);
        launch_axpby(d_e0s, d_e0f, d_e0s, asubv[nsub], 1.0, -1.0, st);
        launch_cluster_correction(D, CL, d_e0s, D.xb, 0, st);
    }
    launch_alpha_combine(D, CL, hybrid, d_alpha_c, d_alpha_tot, st);
    launch_primal(P, nsub, D.xb, d_u, st);
    // KKT report of the solution (GetSolution_Primal_singular_parallel, itersolver.cpp:171-370)
    kkt(d_x, last);
    if (opt.truncate_output) truncate_primal();
    check_launches("primal recovery");
    sync();
    last.time_primal = wall() - tp;
    last.converged = converged ? 1 : 0;
    last.iterations = iters;
    last.final_norm = norm;
    last.tol = tol;
    last.n_lambda = nlam;
    last.n_coarse = ncoarse;
    last.n_subdomains = nsub;
    last.refinement_steps = refine;
    last.sc_bytes = fsc_bytes;
    last.sd_bytes = sd_bytes;
    last.assembly_flops = asm_flops;
    last.time_total = last.time_assembly + last.time_coarse + last.time_pcg + last.time_primal;
    solved = true;
    if (out) *out = last;
}

// KKT residuals of the last primal u (d_u) and dual lambda:
//   kkt_primal = ||K u - f + B1^T lambda (+ B0^T lambda0)|| / ||f - B1^T lambda||
//   kkt_dual   = ||B1 u - c|| / (||c|| + (||c|| < 1 ? 1 : 0))
// (itersolver.cpp:199-251 and :286-353 with only equality constraints).  In
// HYBRID_FETI the corner forces B0^T lambda0 of the cluster problems enter
// the first residual, so that it measures K u = f - B^T lambda for the
// constraints the solution satisfies.
void espreso_feti_s::kkt(const double *lambda, espreso_feti_stats &s) {
    const int64_t nb = D.nbnd;
    double *bt1 = d_kkt, *bt0 = d_kkt + nb, *yl = d_kkt + 2 * nb, *sums = d_kkt + 2 * nb + nlam;
    double *ub = D.yb;  // boundary values of u (scratch)
    launch_boundary_gather(D, lambda, false, bt1, st);
    if (hybrid) {
        ESP_HIP(hipMemsetAsync(bt0, 0, nb * sizeof(double), st));
        launch_b0t(D, CL, bt0, st);
    }
    double *part = d_kkt + 2 * nb + nlam + 2 * (int64_t)nsub + 8;
    launch_kkt_primal(P, nsub, kkt_max_n, d_u, d_f, bt1, hybrid ? bt0 : nullptr, ub, part, sums, st);
    launch_lambda_gather(D, ub, false, yl, st);
    launch_axpby(yl, yl, d_c, nlam, 1.0, -1.0, st);
    launch_dot(yl, yl, nlam, d_partial, sums + 2 * nsub, st);
    launch_dot(d_c, d_c, nlam, d_partial, sums + 2 * nsub + 1, st);
    std::vector<double> h(2 * (size_t)nsub + 2);
    ESP_HIP(hipMemcpyAsync(h.data(), sums, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    sync();
    double num = 0, den = 0;
    for (int q = 0; q < nsub; q++) {
        num += h[2 * q];
        den += h[2 * q + 1];
    }
    num = std::sqrt(num);
    den = std::sqrt(den);
    s.kkt_primal = den > 0 ? num / den : num;
    double nc = std::sqrt(h[2 * nsub + 1]);
    if (std::fabs(nc) < 1) nc += 1.0;
    s.kkt_dual = std::sqrt(h[2 * nsub]) / nc;
}

// output truncation of FETISolver::solve (FETISolver.cpp:131-148), opt-in
void espreso_feti_s::truncate_primal() { launch_truncate(d_u, uoff[nsub], opt.precision, d_partial, st); }

// ------------------------------------------------------ block re-solve ----
// Several load cases on the assembled operators, solved together: one
// projected PCG (itersolver.cpp:1537-1684) per case with its own scalars and
// stop test, advanced in lockstep so that every F and Dirichlet apply reads
// the F_i and S_D bytes once for the whole group of up to 16 cases (block
// SYMV on MFMA).  Each case
// follows FETISolver::solve with its own Matrices::f (FETISolver.cpp:59-149).
void espreso_feti_s::block_setup(int K) {
    if (!SB.items) {
        std::vector<SymvItem> items;
        std::vector<int> item0(nsub), nit(nsub);
        int64_t plen = 0;
        for (int s = 0; s < nsub; s++) {
            const int m = sub[s].m;
            const int nst = (int)strip_count(m, STRIP_F64);
            item0[s] = (int)items.size();
            for (int t = 0; t < nst; t += BLOCK_STRIPS) {
                SymvItem it{};
                it.sub = s;
                it.t0 = t;
                it.t1 = std::min(nst, t + BLOCK_STRIPS);
                it.part_off = plen;
                plen += m - t * STRIP_F64;
                items.push_back(it);
            }
            nit[s] = (int)items.size() - item0[s];
        }
        SB.items = upload(items);
        SB.n_items = (int)items.size();
        SB.item0 = upload(item0);
        SB.nitems = upload(nit);
        SB.plen = plen;
    }
    if (BB.K >= K) return;
    drop_graphs();  // graphs hold the group's buffers
    for (double *q : {BB.x, BB.r, BB.w, BB.z, BB.y, BB.p, BB.Ap, BB.b, BB.e, BB.alpha_c, BB.alpha_tot, BB.e0f, BB.hf,
                      BB.ubf, BB.xbv, BB.yb, BB.f, BB.yf, BB.u, BB.scal, BB.work, BB.e0s, BB.lam0, BB.alphaK,
                      SB.xb, SB.ypart})
        if (q) release(q);
    const int64_t nl = nlam, nb = D.nbnd, nc = std::max(ncoarse, 1), na = std::max<int64_t>(asubv[nsub], 1);
    BB.x = alloc<double>(K * nl);
    BB.r = alloc<double>(K * nl);
    BB.w = alloc<double>(K * nl);
    BB.z = alloc<double>(K * nl);
    BB.y = alloc<double>(K * nl);
    BB.p = alloc<double>(K * nl);
    BB.Ap = alloc<double>(K * nl);
    BB.b = alloc<double>(K * nl);
    BB.e = alloc<double>(K * nc);
    BB.alpha_c = alloc<double>(K * nc);
    BB.alpha_tot = alloc<double>(K * na);
    BB.e0f = alloc<double>(K * na);
    BB.hf = alloc<double>(K * nb);
    BB.ubf = alloc<double>(K * nb);
    BB.xbv = alloc<double>(K * nb);
    BB.yb = alloc<double>(K * nb);
    BB.f = alloc<double>(K * uoff[nsub]);
    BB.yf = alloc<double>(K * std::max<int64_t>(nyf, 1));
    BB.u = alloc<double>(K * uoff[nsub]);
    BB.scal = alloc<double>(8 * K);
    BB.work = alloc<double>(16 * std::max<int64_t>(nyf, 1));
    BB.e0s = alloc<double>(K * n_e0s);
    BB.lam0 = alloc<double>(K * n_lambda0);
    BB.alphaK = alloc<double>(K * na);
    SB.xb = alloc<double>(K * nb);
    SB.ypart = alloc<double>(K * std::max<int64_t>(SB.plen, 1));
    BB.K = K;
}

void espreso_feti_s::lanes_setup() {
    if (!lanes_ready) {
        for (Lane &l : lanes) {
            ESP_HIP(hipStreamCreate(&l.s));
            ESP_HIP(hipEventCreateWithFlags(&l.ev, hipEventDisableTiming));
            l.dsub = alloc<double>(n_dsub);
            l.cd = alloc<double>(std::max(ncoarse, 1));
            l.ce = alloc<double>(std::max(ncoarse, 1));
            l.lambda0 = alloc<double>(n_lambda0);
            l.alphaK = alloc<double>(std::max<int64_t>(asubv[nsub], 1));
            l.e0s = alloc<double>(n_e0s);
            l.partial = alloc<double>(1024);
        }
        ESP_HIP(hipEventCreateWithFlags(&ev_main, hipEventDisableTiming));
        lanes_ready = true;
    }
    // the operators of the current set-up with the lane's own scratch
    for (Lane &l : lanes) {
        l.D = D;
        l.D.dsub = l.dsub;
        l.D.cd = l.cd;
        l.D.ce = l.ce;
        l.C = CL;
        l.C.lambda0 = l.lambda0;
        l.C.alphaK = l.alphaK;
    }
}

// lanes start after the work queued on the main stream; the main stream
// continues after the lanes' work
void espreso_feti_s::fork_lanes() {
    ESP_HIP(hipEventRecord(ev_main, st));
    for (Lane &l : lanes) ESP_HIP(hipStreamWaitEvent(l.s, ev_main, 0));
}

void espreso_feti_s::join_lanes() {
    for (Lane &l : lanes) {
        ESP_HIP(hipEventRecord(l.ev, l.s));
        ESP_HIP(hipStreamWaitEvent(st, l.ev, 0));
    }
}

// hybrid: the cluster correction of the block SYMV output BB.yb for the K
// lambda-space vectors x: F0^-1 and F_i B0^T are read once for the group (main
// stream), the kernel dots and the small alpha solves run per case on the lanes
void espreso_feti_s::cluster_block(int K, const double *x) {
    const int64_t na = std::max<int64_t>(asubv[nsub], 1);
    fork_lanes();
    for (int v = 0; v < K; v++)
        launch_subdomain_kernel_dots(lanes[v % NLANE].D, x + (int64_t)v * nlam, BB.e0s + v * n_e0s,
                                     lanes[v % NLANE].s);
    join_lanes();
    launch_cluster_t_multi(D, CL, K, BB.yb, D.nbnd, BB.lam0, n_lambda0, st);
    fork_lanes();
    for (int v = 0; v < K; v++)
        launch_cluster_alpha(CL, BB.e0s + v * n_e0s, BB.lam0 + v * n_lambda0, BB.alphaK + v * na, lanes[v % NLANE].s);
    join_lanes();
    launch_cluster_apply_multi(D, CL, K, BB.lam0, n_lambda0, BB.alphaK, na, 1, BB.yb, D.nbnd, st);
}

// Y = F X for K lambda-space vectors: one block SYMV on the main stream, the
// cluster correction, then the scatter of each vector on its lane
void espreso_feti_s::apply_F_block(int K, const double *x, double *y) {
    launch_symv_mfma(D, SB, false, K, x, false, BB.yb, st);
    if (hybrid) cluster_block(K, x);
    fork_lanes();
    for (int v = 0; v < K; v++)
        launch_lambda_gather(lanes[v % NLANE].D, BB.yb + v * D.nbnd, false, y + (int64_t)v * nlam, lanes[v % NLANE].s);
    join_lanes();
}

// one PCG iteration of every active case (the block applies cover all K);
// the per-case chains between the block applies run on the lanes
void espreso_feti_s::step_block(int K, const std::vector<char> &active, bool first) {
    const int64_t nl = nlam;
    auto sc = [&](int v, int q) { return BB.scal + 8 * v + q; };  // 0 yw, 1 yw_prev, 2 pAp, 3 ww
    auto lane = [&](int v) -> Lane & { return lanes[v % NLANE]; };
    fork_lanes();
    for (int v = 0; v < K; v++)
        if (active[v]) launch_project(lane(v).D, BB.r + v * nl, BB.w + v * nl, 0, lane(v).s);
    join_lanes();
    const bool dir = opt.preconditioner == ESPRESO_PREC_DIRICHLET;
    if (dir) {
        launch_symv_mfma(D, SB, true, K, BB.w, true, BB.yb, st);
    } else if (opt.preconditioner != ESPRESO_PREC_NONE) {
        for (int v = 0; v < K; v++)
            if (active[v]) apply_prec(BB.w + v * nl, BB.z + v * nl);  // main stream (shared boundary scratch)
    }
    fork_lanes();
    for (int v = 0; v < K; v++) {
        if (!active[v]) continue;
        Lane &L = lane(v);
        if (opt.preconditioner == ESPRESO_PREC_NONE) {
            ESP_HIP(hipMemcpyAsync(BB.y + v * nl, BB.w + v * nl, nl * sizeof(double), hipMemcpyDeviceToDevice, L.s));
        } else {
            if (dir) launch_lambda_gather(L.D, BB.yb + v * D.nbnd, true, BB.z + v * nl, L.s);
            launch_project(L.D, BB.z + v * nl, BB.y + v * nl, 0, L.s);
        }
        if (!first) launch_copy_scalar(sc(v, 0), sc(v, 1), L.s);
        launch_dot(BB.y + v * nl, BB.w + v * nl, nl, L.partial, sc(v, 0), L.s);
        launch_p_update(BB.p + v * nl, BB.y + v * nl, nl, sc(v, 0), sc(v, 1), first, L.s);
    }
    join_lanes();
    launch_symv_mfma(D, SB, false, K, BB.p, false, BB.yb, st);
    if (hybrid) cluster_block(K, BB.p);
    fork_lanes();
    for (int v = 0; v < K; v++) {
        if (!active[v]) continue;
        Lane &L = lane(v);
        double *ybv = BB.yb + v * D.nbnd, *Apv = BB.Ap + v * nl;
        launch_lambda_gather(L.D, ybv, false, Apv, L.s);
        launch_dot(BB.p + v * nl, Apv, nl, L.partial, sc(v, 2), L.s);
        launch_xr_update(BB.x + v * nl, BB.r + v * nl, BB.p + v * nl, Apv, nl, sc(v, 0), sc(v, 2), L.s);
        launch_dot(BB.w + v * nl, BB.w + v * nl, nl, L.partial, sc(v, 3), L.s);
    }
    join_lanes();
}

void espreso_feti_s::step_block_graph(int K) {
    for (const BlockGraph &g : block_graphs)
        if (g.K == K) {
            ESP_HIP(hipGraphLaunch(g.exec, st));
            return;
        }
    const std::vector<char> all(K, 1);
    hipGraph_t graph = nullptr;
    ESP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        step_block(K, all, false);
    } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    ESP_HIP(hipStreamEndCapture(st, &graph));
    BlockGraph g{K, nullptr};
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    ESP_HIP(e);
    block_graphs.push_back(g);
    ESP_HIP(hipGraphLaunch(g.exec, st));
}

void espreso_feti_s::resolve_group(int K, const double *f, double *u, int32_t *iters, double *norms,
                                   espreso_feti_stats &tot) {
    const double t0 = wall();
    block_setup(K);
    lanes_setup();
    const int64_t nl = nlam, nb = D.nbnd, nt = uoff[nsub], na = std::max<int64_t>(asubv[nsub], 1);
    const int64_t nc = std::max(ncoarse, 1);
    // loads: e0f, e on the host; f, y_f, hf (one band-solve launch) and b on the device
    hvec<double> e0f;
    std::vector<double> e;
    ESP_HIP(hipMemcpyAsync(BB.f, f, K * nt * sizeof(double), hipMemcpyHostToDevice, st));
    for (int v = 0; v < K; v++) {
        rhs_host(f + v * nt, e0f, e);
        if (!e0f.empty())
            ESP_HIP(hipMemcpyAsync(BB.e0f + v * na, e0f.data(), e0f.size() * sizeof(double), hipMemcpyHostToDevice, st));
        if (!e.empty())
            ESP_HIP(hipMemcpyAsync(BB.e + v * nc, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, st));
        sync();  // host vectors e0f, e are reused by the next case
    }
    launch_rhs_prepare_multi(P, nsub, K, BB.f, nt, BB.hf, nb, BB.yf, nyf, BB.work, st);
    for (int v = 0; v < K; v++) compute_b_into(BB.hf + v * nb, BB.ubf + v * nb, BB.e0f + v * na, BB.b + v * nl);
    // lambda0 = G^T (GG^T)^-1 e ; r = b - F lambda0 ; tol = precision * min(||P r0||, ||b||)
    for (int v = 0; v < K; v++) launch_project(D, BB.e + v * nc, BB.x + v * nl, 1, st);
    apply_F_block(K, BB.x, BB.Ap);
    for (int v = 0; v < K; v++) {
        launch_axpby(BB.r + v * nl, BB.b + v * nl, BB.Ap + v * nl, nl, 1.0, -1.0, st);
        launch_project(D, BB.r + v * nl, BB.w + v * nl, 0, st);
        launch_dot(BB.w + v * nl, BB.w + v * nl, nl, d_partial, BB.scal + 8 * v + 4, st);
        launch_dot(BB.b + v * nl, BB.b + v * nl, nl, d_partial, BB.scal + 8 * v + 5, st);
    }
    std::vector<double> hs(8 * K);
    ESP_HIP(hipMemcpyAsync(hs.data(), BB.scal, hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    sync();
    const double t1 = wall();
    std::vector<double> tol(K), nrm(K, 0.0);
    std::vector<char> active(K, 1);
    std::vector<int> it_v(K, 0);
    for (int v = 0; v < K; v++) tol[v] = opt.precision * std::min(std::sqrt(hs[8 * v + 4]), std::sqrt(hs[8 * v + 5]));
    bool all_active = true;
    for (int it = 0; it < opt.max_iterations; it++) {
        if (it > 0 && all_active && use_graphs) step_block_graph(K);
        else step_block(K, active, it == 0);
        ESP_HIP(hipMemcpyAsync(hs.data(), BB.scal, hs.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        bool any = false;
        for (int v = 0; v < K; v++) {
            if (!active[v]) continue;
            ESP_CHECK(std::isfinite(hs[8 * v + 3]), ESPRESO_ERR_NUMERIC, "block PCG produced a non-finite residual");
            nrm[v] = std::sqrt(hs[8 * v + 3]);
            it_v[v] = it + 1;
            if (nrm[v] < tol[v]) active[v] = 0;
            any |= active[v] != 0;
            all_active = all_active && active[v];
        }
        if (!any) break;
    }
    sync();
    const double t2 = wall();
    // primal per case: u = K~^+ (f - B^T lambda) + R~ alpha
    for (int v = 0; v < K; v++) launch_project(D, BB.r + v * nl, BB.alpha_c + v * nc, 2, st);
    launch_symv_mfma(D, SB, false, K, BB.x, false, BB.yb, st);
    for (int v = 0; v < K; v++) {
        double *xb = BB.xbv + v * nb;
        launch_axpby(xb, BB.ubf + v * nb, BB.yb + v * nb, nb, 1.0, -1.0, st);
        if (hybrid) {
            launch_subdomain_kernel_dots(D, BB.x + v * nl, d_e0s, st);
            launch_axpby(d_e0s, BB.e0f + v * na, d_e0s, asubv[nsub], 1.0, -1.0, st);
            launch_cluster_correction(D, CL, d_e0s, xb, 0, st);
        }
        launch_alpha_combine(D, CL, hybrid, BB.alpha_c + v * nc, BB.alpha_tot + v * na, st);
    }
    launch_primal_multi(P, nsub, K, BB.xbv, nb, BB.yf, nyf, BB.alpha_tot, na, BB.u, nt, BB.work, st);
    ESP_HIP(hipMemcpyAsync(u, BB.u, K * nt * sizeof(double), hipMemcpyDeviceToHost, st));
    sync();
    for (int v = 0; v < K; v++) {
        if (iters) iters[v] = it_v[v];
        if (norms) norms[v] = nrm[v];
        tot.iterations = std::max(tot.iterations, it_v[v]);
        tot.final_norm = std::max(tot.final_norm, nrm[v]);
        if (!(nrm[v] < tol[v])) tot.converged = 0;
    }
    check_launches("block re-solve");
    tot.time_coarse += t1 - t0;
    tot.time_pcg += t2 - t1;
    tot.time_primal += wall() - t2;
}

bool espreso_feti_s::resolve_block(int nrhs, const double *f, double *u, int32_t *iters, double *norms,
                                   espreso_feti_stats *stats) {
    ESP_CHECK(ready, ESPRESO_ERR_STATE, "espreso_feti_setup not called");
    const int64_t nt = uoff[nsub];
    espreso_feti_stats tot{};
    tot.n_lambda = nlam;
    tot.n_coarse = ncoarse;
    tot.n_subdomains = nsub;
    tot.kkt_primal = tot.kkt_dual = -1.0;  // not computed for load cases
    tot.converged = 1;
    if (opt.sc_precision != 0 || D.sd_prec != 0) {
        // low-precision storage: the cases one after another (refinement per
        // case); the instance's own solution (lambda, u, stats) is restored after
        std::vector<std::vector<double>> keep(nsub);
        for (int s = 0; s < nsub; s++) keep[s] = sub[s].f;
        double *save_x = alloc<double>(nlam), *save_u = alloc<double>(nt);
        ESP_HIP(hipMemcpyAsync(save_x, d_x, nlam * sizeof(double), hipMemcpyDeviceToDevice, st));
        ESP_HIP(hipMemcpyAsync(save_u, d_u, nt * sizeof(double), hipMemcpyDeviceToDevice, st));
        const espreso_feti_stats save_last = last;
        const bool save_solved = solved;
        for (int v = 0; v < nrhs; v++) {
            for (int s = 0; s < nsub; s++)
                std::copy(f + v * nt + uoff[s], f + v * nt + uoff[s] + sub[s].n, sub[s].f.begin());
            resolve_rhs();
            espreso_feti_stats sv{};
            solve(&sv);
            ESP_HIP(hipMemcpy(u + v * nt, d_u, nt * sizeof(double), hipMemcpyDeviceToHost));
            if (iters) iters[v] = sv.iterations;
            if (norms) norms[v] = sv.final_norm;
            tot.iterations = std::max(tot.iterations, sv.iterations);
            tot.final_norm = std::max(tot.final_norm, sv.final_norm);
            tot.time_pcg += sv.time_pcg;
            tot.time_primal += sv.time_primal;
            tot.converged &= sv.converged;
        }
        for (int s = 0; s < nsub; s++) sub[s].f = keep[s];
        resolve_rhs();
        ESP_HIP(hipMemcpyAsync(d_x, save_x, nlam * sizeof(double), hipMemcpyDeviceToDevice, st));
        ESP_HIP(hipMemcpyAsync(d_u, save_u, nt * sizeof(double), hipMemcpyDeviceToDevice, st));
        sync();
        release(save_x);
        release(save_u);
        last = save_last;
        solved = save_solved;
        tot.time_total = tot.time_pcg + tot.time_primal;
        if (stats) *stats = tot;
        return tot.converged != 0;
    }
    for (int done = 0; done < nrhs;) {
        const int rem = nrhs - done;
        const int K = std::min(rem, 16);
        resolve_group(K, f + done * nt, u + done * nt, iters ? iters + done : nullptr, norms ? norms + done : nullptr,
                      tot);
        done += K;
    }
    tot.time_total = tot.time_coarse + tot.time_pcg + tot.time_primal;
    if (stats) *stats = tot;
    return tot.converged != 0;
}

// --------------------------------------------------- coarse space layout ----
// GG^T (dense): sum over lambda of the outer products of G columns, G = -(B Rt)^T
void espreso_feti_s::build_ggt(const double *Rt, const std::vector<int64_t> &rtoff, const std::vector<int> &ck,
                               const std::vector<int> &cblk, const std::vector<int> &coff, hvec<double> &GGt) {
    const int nc = ncoarse;
    const std::vector<int64_t> &boff = h_boff;
    int nblk = 0;
    for (int s = 0; s < nsub; s++) nblk = std::max(nblk, cblk[s] + 1);
    std::vector<std::vector<int>> subs_of(nblk);
    for (int s = 0; s < nsub; s++) subs_of[cblk[s]].push_back(s);
    // each thread owns the columns of its coarse blocks (zeroes and fills
    // them): no shared writes, and a fixed order of the sums
    GGt.resize((size_t)nc * nc);
#pragma omp parallel for schedule(dynamic)
    for (int blk = 0; blk < nblk; blk++) {
        if (subs_of[blk].empty()) continue;
        const int o2 = coff[blk], k2 = ck[subs_of[blk][0]];
        std::fill(GGt.begin() + (size_t)o2 * nc, GGt.begin() + (size_t)(o2 + k2) * nc, 0.0);
        for (const int s2 : subs_of[blk]) {
            const SubHost &h2 = sub[s2];
            for (size_t q = 0; q < h2.lam.size(); q++) {
                const int l = h2.lam[q];
                const int e2 = le_sub_h[2 * l] == s2 ? 0 : 1;
                const int64_t j2 = le_b_h[2 * l + e2] - boff[s2];
                for (int a = 0; a < le_cnt_h[l]; a++) {
                    const int s1 = le_sub_h[2 * l + a];
                    const int64_t j1 = le_b_h[2 * l + a] - boff[s1];
                    const double g = le_sign_h[2 * l + a] * le_sign_h[2 * l + e2];
                    const int m1 = sub[s1].m, m2 = h2.m;
                    const int o1 = coff[cblk[s1]];
                    for (int j = 0; j < ck[s2]; j++) {
                        const double r2 = g * Rt[rtoff[s2] + j2 + (int64_t)j * m2];
                        double *col = GGt.data() + (size_t)(o2 + j) * nc + o1;
                        for (int i = 0; i < ck[s1]; i++) col[i] += Rt[rtoff[s1] + j1 + (int64_t)i * m1] * r2;
                    }
                }
            }
        }
    }
}

// Total FETI: one coarse block per subdomain, basis R_b, N_s = I
void espreso_feti_s::coarse_layout_subdomains() {
    std::vector<int> cblk(nsub), ck(nsub), coff(nsub), bptr(nsub + 1), bsub(nsub);
    std::vector<int64_t> noff(nsub + 1, 0);
    ncoarse = 0;
    for (int s = 0; s < nsub; s++) {
        cblk[s] = s;
        ck[s] = sub[s].kdim;
        coff[s] = ncoarse;
        ncoarse += sub[s].kdim;
        bptr[s] = s;
        bsub[s] = s;
        noff[s + 1] = noff[s] + (int64_t)sub[s].kdim * sub[s].kdim;
    }
    bptr[nsub] = nsub;
    std::vector<double> N(noff[nsub], 0.0);
    for (int s = 0; s < nsub; s++)
        for (int a = 0; a < sub[s].kdim; a++) N[noff[s] + a + (int64_t)a * sub[s].kdim] = 1.0;
    D.nblk = nsub;
    D.ncoarse = ncoarse;
    D.cblk = upload(cblk);
    D.ck = upload(ck);
    D.coff = upload(coff);
    D.blk_ptr = upload(bptr);
    D.blk_sub = upload(bsub);
    D.Rt = D.Rb;
    D.rtoff = D.rboff;
    D.dsub_off = D.asub;
    n_dsub = std::max<int64_t>(1, asubv[nsub]);
    D.dsub = alloc<double>(n_dsub);
    D.N = upload(N);
    D.noff = upload(noff);
    D.cd = alloc<double>(ncoarse);
    D.ce = alloc<double>(ncoarse);
    d_alpha_c = alloc<double>(ncoarse);
    d_e = upload(e0f_h);
    d_GGt = alloc<double>((size_t)ncoarse * ncoarse);
    if (!ggt_st) ESP_HIP(hipStreamCreate(&ggt_st));
    if (!ggt_ev) ESP_HIP(hipEventCreateWithFlags(&ggt_ev, hipEventDisableTiming));
    ESP_HIP(hipEventRecord(ggt_ev, st));  // behind the allocation
    ESP_HIP(hipStreamWaitEvent(ggt_st, ggt_ev, 0));
    ggt_thread = std::thread([this, ck, cblk, coff] {
        try {
            ESP_HIP(hipSetDevice(opt.device));
            hvec<double> GGt;
            build_ggt(Rb_h.data(), rboff_h, ck, cblk, coff, GGt);
            ESP_HIP(hipMemcpyAsync(d_GGt, GGt.data(), GGt.size() * sizeof(double), hipMemcpyHostToDevice, ggt_st));
            ESP_HIP(hipStreamSynchronize(ggt_st));
        } catch (...) {
            ggt_err = std::current_exception();
        }
    });
}

// Hybrid FETI, host part (once per load; cluster.cpp:1186-1291, 2018-2159):
// B0 rows per cluster and corner entries per subdomain, G0 = -(B0 R)^T, the
// cluster kernel N = ker G0^T (Jacobi on G0 G0^T), the coarse space with one
// block per cluster and basis R~_s = R_b,s N_s, GG^T and e.
void espreso_feti_s::load_clusters() {
    const std::vector<int64_t> &boff = h_boff;
    n0_h.assign(ncl, 0);
    kt_h.assign(ncl, 0);
    std::vector<int> kn(ncl, 0), sfirst(ncl), scount(ncl);
    std::vector<int64_t> rowoff(ncl + 1, 0), eoff(ncl, 0);
    for (int c = 0; c < ncl; c++) {
        const auto &L = cl_subs[c];
        for (size_t q = 1; q < L.size(); q++)
            ESP_CHECK(L[q] == L[q - 1] + 1, ESPRESO_ERR_ARG, "HYBRID_FETI: subdomains of a cluster must be consecutive");
        int mx = -1;
        for (int s : L)
            for (int r : sub[s].b0row) mx = std::max(mx, r);
        n0_h[c] = mx + 1;
        rowoff[c + 1] = rowoff[c] + n0_h[c];
        for (int s : L) kt_h[c] += sub[s].kdim;
        eoff[c] = asubv[L[0]];
        sfirst[c] = L[0];
        scount[c] = (int)L.size();
    }
    // slots: the B0 rows touching a subdomain, ascending, each with its
    // entries (boundary index, coefficient) in input order; and the entries of
    // every cluster row over all its subdomains (g0 = B0 y_b)
    nq_h.assign(nsub, 0);
    qoff_h.assign(nsub + 1, 0);
    std::vector<int64_t> fb0off(nsub + 1, 0);
    fccoff_h.assign(nsub + 1, 0);
    qrow_h.clear();
    std::vector<int64_t> qe_ptr(1, 0);
    std::vector<int> qe_j;
    std::vector<double> qe_c;
    std::vector<std::vector<std::pair<int64_t, double>>> rowent(rowoff[ncl]);
    max_qe_h = 0;
    for (int s = 0; s < nsub; s++) {
        const SubHost &h = sub[s];
        const int c = h.cluster;
        std::map<int, std::vector<std::pair<int, double>>> slots;
        for (size_t q = 0; q < h.b0row.size(); q++) {
            const int r = h.b0row[q];
            ESP_CHECK(r >= 0 && r < n0_h[c], ESPRESO_ERR_ARG, "B0 row out of range");
            slots[r].push_back({h.iperm[h.b0dof[q]] - h.ni, h.b0sign[q]});
        }
        nq_h[s] = (int)slots.size();
        ESP_CHECK(nq_h[s] <= 256, ESPRESO_ERR_ARG, "HYBRID_FETI: more than 256 B0 rows touch one subdomain");
        qoff_h[s + 1] = qoff_h[s] + nq_h[s];
        fb0off[s + 1] = fb0off[s] + (int64_t)h.m * nq_h[s];
        fccoff_h[s + 1] = fccoff_h[s] + (int64_t)nq_h[s] * nq_h[s];
        for (const auto &kv : slots) {
            const int64_t row = rowoff[c] + kv.first;
            qrow_h.push_back((int)row);
            for (const auto &e : kv.second) {
                qe_j.push_back(e.first);
                qe_c.push_back(e.second);
                rowent[row].push_back({boff[s] + e.first, e.second});
            }
            qe_ptr.push_back((int64_t)qe_j.size());
            max_qe_h = std::max<int64_t>(max_qe_h, (int64_t)kv.second.size());
        }
    }
    std::vector<int64_t> re_ptr(rowoff[ncl] + 1, 0), re_b;
    std::vector<double> re_c;
    for (int64_t r = 0; r < rowoff[ncl]; r++) {
        ESP_CHECK(!rowent[r].empty(), ESPRESO_ERR_ARG, "HYBRID_FETI: a B0 row without entries");
        for (const auto &e : rowent[r]) {
            re_b.push_back(e.first);
            re_c.push_back(e.second);
        }
        re_ptr[r + 1] = (int64_t)re_b.size();
    }

    // G0^T (n0 x kt per cluster) and the cluster kernel
    std::vector<int64_t> g0off(ncl + 1, 0), ncoff(ncl + 1, 0);
    f0off_h.assign(ncl + 1, 0);
    saoff_h.assign(ncl + 1, 0);
    for (int c = 0; c < ncl; c++) {
        g0off[c + 1] = g0off[c] + (int64_t)n0_h[c] * kt_h[c];
        f0off_h[c + 1] = f0off_h[c] + (int64_t)n0_h[c] * n0_h[c];
        saoff_h[c + 1] = saoff_h[c] + (int64_t)kt_h[c] * kt_h[c];
    }
    std::vector<double> G0T(std::max<int64_t>(1, g0off[ncl]), 0.0);
    std::vector<std::vector<double>> Nc(ncl);
    std::string err;
#pragma omp parallel for schedule(dynamic)
    for (int c = 0; c < ncl; c++) {
        const int n = n0_h[c], k = kt_h[c];
        double *G = G0T.data() + g0off[c];
        int aoff = 0;
        for (int s : cl_subs[c]) {
            const SubHost &h = sub[s];
            for (int64_t q = qoff_h[s]; q < qoff_h[s + 1]; q++) {
                const int r = qrow_h[q] - (int)rowoff[c];
                for (int64_t e = qe_ptr[q]; e < qe_ptr[q + 1]; e++)
                    for (int a = 0; a < h.kdim; a++)
                        G[r + (int64_t)(aoff + a) * n] -= qe_c[e] * Rb_h[rboff_h[s] + qe_j[e] + (int64_t)a * h.m];
            }
            aoff += h.kdim;
        }
        std::vector<double> GGc((size_t)k * k, 0.0), w, V;
        for (int a = 0; a < k; a++)
            for (int b = 0; b < k; b++) {
                double v = 0;
                for (int r = 0; r < n; r++) v += G[r + (int64_t)a * n] * G[r + (int64_t)b * n];
                GGc[a + (size_t)b * k] = v;
            }
        jacobi_eigen(k, GGc, w, V);
        double wmax = 0;
        for (double x : w) wmax = std::max(wmax, std::fabs(x));
        int kk = 0;
        for (int i = 0; i < k; i++)
            if (wmax == 0.0 || std::fabs(w[i]) <= 1e-10 * wmax) {
                Nc[c].insert(Nc[c].end(), V.begin() + (size_t)i * k, V.begin() + (size_t)(i + 1) * k);
                kk++;
            }
        kn[c] = kk;
        if (kk == 0 && k > 0) {
#pragma omp critical
            err = "HYBRID_FETI: cluster " + std::to_string(c) + " has no rigid modes left (over-constrained B0)";
        }
    }
    ESP_CHECK(err.empty(), ESPRESO_ERR_NUMERIC, err);
    for (int c = 0; c < ncl; c++) {
        ESP_CHECK(kn[c] <= 64, ESPRESO_ERR_NUMERIC, "HYBRID_FETI: cluster kernel dimension above 64");
        ncoff[c + 1] = ncoff[c] + (int64_t)kt_h[c] * kn[c];
    }
    std::vector<double> Nall(std::max<int64_t>(1, ncoff[ncl]));
    for (int c = 0; c < ncl; c++) std::copy(Nc[c].begin(), Nc[c].end(), Nall.begin() + ncoff[c]);

    CL.ncl = ncl;
    CL.max_n0 = *std::max_element(n0_h.begin(), n0_h.end());
    ESP_CHECK(CL.max_n0 <= 8192, ESPRESO_ERR_ARG, "HYBRID_FETI: more than 8192 B0 rows in one cluster");
    CL.max_kt = *std::max_element(kt_h.begin(), kt_h.end());
    CL.max_kn = *std::max_element(kn.begin(), kn.end());
    CL.n0 = upload(n0_h);
    CL.kt = upload(kt_h);
    CL.kn = upload(kn);
    CL.sfirst = upload(sfirst);
    CL.scount = upload(scount);
    CL.rowoff = upload(rowoff);
    CL.eoff = upload(eoff);
    CL.f0off = upload(f0off_h);
    CL.g0off = upload(g0off);
    CL.saoff = upload(saoff_h);
    CL.ncoff = upload(ncoff);
    CL.re_ptr = upload(re_ptr);
    CL.re_b = upload(re_b);
    CL.re_c = upload(re_c);
    CL.qe_ptr = upload(qe_ptr);
    CL.qe_j = upload(qe_j);
    CL.qe_c = upload(qe_c);
    CL.G0T = upload(G0T);
    CL.H = alloc<double>(std::max<int64_t>(1, g0off[ncl]));
    CL.Nc = upload(Nall);
    CL.F0inv = alloc<double>(std::max<int64_t>(1, f0off_h[ncl]));
    CL.Sainv = alloc<double>(std::max<int64_t>(1, saoff_h[ncl]));
    CL.nq = upload(nq_h);
    CL.qoff = upload(qoff_h);
    CL.qrow = upload(qrow_h);
    CL.fb0off = upload(fb0off);
    CL.fccoff = upload(fccoff_h);
    CL.asub = D.asub;
    CL.FB0 = alloc<double>(fb0off[nsub]);
    if (opt.sc_precision) CL.FB0f = alloc<float>(fb0off[nsub]);
    CL.Fcc = alloc<double>(std::max<int64_t>(1, fccoff_h[nsub]));
    n_lambda0 = std::max<int64_t>(1, rowoff[ncl]);
    CL.lambda0 = alloc<double>(n_lambda0);
    CL.alphaK = alloc<double>(asubv[nsub]);
    CL.rho_s = alloc<double>(nsub);
    d_Fcc = CL.Fcc;

    // coarse space: one block per cluster, basis Rt_s = R_b,s N_s
    std::vector<int> cblk(nsub), ck(nsub), coff(ncl), bptr(ncl + 1, 0), bsub;
    std::vector<int64_t> rtoff(nsub + 1, 0), noff(nsub + 1, 0), dsub_off(nsub + 1, 0);
    ncoarse = 0;
    for (int c = 0; c < ncl; c++) {
        coff[c] = ncoarse;
        ncoarse += kn[c];
        for (int s : cl_subs[c]) bsub.push_back(s);
        bptr[c + 1] = (int)bsub.size();
    }
    for (int s = 0; s < nsub; s++) {
        const int c = sub[s].cluster;
        cblk[s] = c;
        ck[s] = kn[c];
        rtoff[s + 1] = rtoff[s] + (int64_t)sub[s].m * kn[c];
        noff[s + 1] = noff[s] + (int64_t)sub[s].kdim * kn[c];
        dsub_off[s + 1] = dsub_off[s] + kn[c];
    }
    std::vector<double> Rt(std::max<int64_t>(1, rtoff[nsub]), 0.0), Ns(std::max<int64_t>(1, noff[nsub]), 0.0),
        e(std::max(ncoarse, 1), 0.0);
    for (int c = 0; c < ncl; c++) {
        const int k = kt_h[c];
        int aoff = 0;
        for (int s : cl_subs[c]) {
            const SubHost &h = sub[s];
            for (int b = 0; b < kn[c]; b++) {
                for (int a = 0; a < h.kdim; a++) {
                    const double nab = Nc[c][(aoff + a) + (size_t)b * k];
                    Ns[noff[s] + a + (int64_t)b * h.kdim] = nab;
                    e[coff[c] + b] += nab * e0f_h[asubv[s] + a];
                    for (int j = 0; j < h.m; j++)
                        Rt[rtoff[s] + j + (int64_t)b * h.m] += Rb_h[rboff_h[s] + j + (int64_t)a * h.m] * nab;
                }
            }
            aoff += h.kdim;
        }
    }
    ESP_CHECK(ncoarse > 0, ESPRESO_ERR_NUMERIC, "HYBRID_FETI: empty coarse space");
    Ns_h = Ns;
    noff_h = noff;
    ccoff_h = coff;
    cblk_h = cblk;
    D.nblk = ncl;
    D.ncoarse = ncoarse;
    D.cblk = upload(cblk);
    D.ck = upload(ck);
    D.coff = upload(coff);
    D.blk_ptr = upload(bptr);
    D.blk_sub = upload(bsub);
    D.rtoff = upload(rtoff);
    D.Rt = upload(Rt);
    D.N = upload(Ns);
    D.noff = upload(noff);
    D.dsub_off = upload(dsub_off);
    n_dsub = std::max<int64_t>(1, dsub_off[nsub]);
    D.dsub = alloc<double>(n_dsub);
    D.cd = alloc<double>(ncoarse);
    D.ce = alloc<double>(ncoarse);
    d_alpha_c = alloc<double>(ncoarse);
    hvec<double> GGt;
    build_ggt(Rt.data(), rtoff, ck, cblk, coff, GGt);
    d_GGt = upload(GGt);
    d_e = upload(e);
}

// In-place inverses of nb symmetric positive definite matrices pool + off[i]
// (sizes n[i], full storage) with the MFMA Cholesky / forward-substitution /
// W^T W kernels; sizes 0 are skipped.
void espreso_feti_s::invert_spd_batched(double *pool, const std::vector<int64_t> &off, const std::vector<int> &n,
                                        const char *what) {
    std::vector<int> idx;
    int64_t wsz = 0, dsz = 0;
    int max_n = 0, max_nt = 0;
    for (size_t i = 0; i < n.size(); i++)
        if (n[i] > 0) {
            idx.push_back((int)i);
            const int nt = (n[i] + TILE - 1) / TILE;
            wsz += (int64_t)n[i] * n[i];
            dsz += (int64_t)nt * TILE * TILE;
            max_n = std::max(max_n, n[i]);
            max_nt = std::max(max_nt, nt);
        }
    if (idx.empty()) return;
    const int nb = (int)idx.size();
    double *W = alloc<double>(wsz), *Dv = alloc<double>(dsz);
    int *info = alloc<int>(nb);
    ESP_HIP(hipMemsetAsync(info, 0, nb * sizeof(int), st));
    std::vector<DenseDesc> desc(nb);
    int64_t wo = 0, dofs = 0;
    for (int b = 0; b < nb; b++) {
        const int i = idx[b];
        DenseDesc &d = desc[b];
        d = DenseDesc{};
        d.S = pool + off[i];
        d.W = W + wo;
        d.Dinv = Dv + dofs;
        d.info = info + b;
        d.ni = 0;
        d.m = n[i];
        d.ntm = d.ntc = (n[i] + TILE - 1) / TILE;
        wo += (int64_t)n[i] * n[i];
        dofs += (int64_t)d.ntm * TILE * TILE;
    }
    DenseDesc *dd = upload(desc);
    // few matrices: W = L^-1 by recursive doubling (the forward substitution
    // would be one short launch per tile row)
    launch_potrf_rl(dd, nb, max_nt, st);
    launch_trtri(dd, nb, max_nt, st);
    launch_syrk(dd, nb, 1, max_nt, st);
    launch_symmetrize_desc(dd, nb, max_n, st);
    std::vector<int> hinfo(nb);
    ESP_HIP(hipMemcpyAsync(hinfo.data(), info, nb * sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    for (int b = 0; b < nb; b++)
        ESP_CHECK(hinfo[b] == 0, ESPRESO_ERR_NUMERIC,
                  std::string("HYBRID_FETI: ") + what + " of cluster " + std::to_string(idx[b]) +
                      " is not positive definite");
    release(W);
    release(Dv);
    release(info);
    release(dd);
}

// Hybrid FETI, device part (every setup; cluster.cpp:1293-2016): F0 from the
// corner blocks of the assembly, F0^-1, H = F0^-1 G0^T, Sa = G0 H regularised
// on the cluster kernel, Sa^-1.
void espreso_feti_s::cluster_setup() {
    launch_cluster_setup_f0(CL, st);
    invert_spd_batched(CL.F0inv, f0off_h, n0_h, "F0");
    launch_cluster_setup_sa(CL, CL.Sainv, st);
    invert_spd_batched(CL.Sainv, saoff_h, kt_h, "regularised Sa");
}

// ================================================================ C ABI ====
#define ESP_API_BEGIN try {
#define ESP_API_END                      \
    }                                    \
    catch (const Error &e) {             \
        g_last_error = e.what();         \
        return e.code;                   \
    }                                    \
    catch (const std::exception &e) {    \
        g_last_error = e.what();         \
        return ESPRESO_ERR_ARG;          \
    }                                    \
    return ESPRESO_OK;

extern "C" {

const char *espreso_last_error(void) { return g_last_error.c_str(); }
int espreso_version(void) { return 1; }


int espreso_feti_create(espreso_feti **h, int32_t n_subdomains, const espreso_feti_options *o) {
    ESP_API_BEGIN
    ESP_CHECK(h && o && n_subdomains > 0, ESPRESO_ERR_ARG, "invalid arguments to espreso_feti_create");
    *h = nullptr;
    int ndev = 0;
    ESP_HIP(hipGetDeviceCount(&ndev));
    ESP_CHECK(ndev > 0, ESPRESO_ERR_HIP, "no HIP device");
    ESP_CHECK(o->device >= 0 && o->device < ndev, ESPRESO_ERR_ARG, "device ordinal out of range");
    ESP_HIP(hipSetDevice(o->device));
    keep_pooled_memory(o->device);
    auto f = std::make_unique<espreso_feti_s>();
    f->opt = *o;
    if (f->opt.max_iterations <= 0) f->opt.max_iterations = 200;
    if (!(f->opt.precision > 0)) f->opt.precision = 1e-5;
    f->nsub = n_subdomains;
    f->sub.resize(n_subdomains);
    ESP_HIP(hipStreamCreate(&f->st));
    *h = f.release();
    ESP_API_END
}

int espreso_feti_destroy(espreso_feti *h) {
    ESP_API_BEGIN
    if (h) ESP_HIP(hipSetDevice(h->opt.device));
    delete h;
    ESP_API_END
}

int espreso_feti_sizes(espreso_feti *h, int32_t *sizes) {
    ESP_API_BEGIN
    ESP_CHECK(h && sizes, ESPRESO_ERR_ARG, "null argument");
    sizes[0] = h->nsub;
    sizes[1] = h->nlam;
    sizes[2] = h->ncoarse;
    ESP_API_END
}

int espreso_feti_set_subdomain(espreso_feti *h, int32_t i, int32_t n, const int32_t *rowptr, const int32_t *colidx,
                               const double *values, const double *f, int32_t kdim, const double *R, int32_t nlam,
                               const int32_t *lambda, const int32_t *dof, const double *sign, const double *scale) {
    ESP_API_BEGIN
    ESP_CHECK(h && i >= 0 && i < h->nsub && n > 0, ESPRESO_ERR_ARG, "invalid subdomain");
    ESP_CHECK(kdim >= 1 && kdim <= 6, ESPRESO_ERR_ARG, "kernel dimension must be 1..6");
    ESP_CHECK(nlam > 0, ESPRESO_ERR_ARG, "subdomain without multipliers (not floating-glued)");
    ESP_CHECK(!h->loaded, ESPRESO_ERR_STATE,
              "setup already done (new K values: espreso_feti_update_values; new loads: espreso_feti_set_rhs)");
    SubHost &s = h->sub[i];
    s.n = n;
    s.rowptr.assign(rowptr, rowptr + n + 1);
    ESP_CHECK(s.rowptr[0] == 0, ESPRESO_ERR_ARG, "CSR must be 0-based");
    const int nnz = s.rowptr[n];
    // subdomains of at most 65536 DOFs keep 16-bit column indices in pinned
    // memory (read once here for the range check, then uploaded from there)
    uint16_t *c16 = n <= 65536 ? h->pin_alloc<uint16_t>(h->pin_idx, (size_t)nnz, s.pin_pos) : nullptr;
    bool sorted = true;  // strictly increasing columns in every row
    int64_t nlow = 0;    // entries with column <= row
    if (c16) {
        std::vector<int>().swap(s.colidx);
        for (int r = 0; r < n; r++) {
            int prev = -1;
            for (int q = s.rowptr[r]; q < s.rowptr[r + 1]; q++) {
                const int c = colidx[q];
                ESP_CHECK(c >= 0 && c < n, ESPRESO_ERR_ARG, "CSR column out of range");
                c16[q] = (uint16_t)c;
                sorted = sorted && c > prev;
                prev = c;
                nlow += c <= r;
            }
        }
    } else {
        s.colidx.assign(colidx, colidx + nnz);
        for (int q = 0; q < nnz; q++)
            ESP_CHECK(s.colidx[q] >= 0 && s.colidx[q] < n, ESPRESO_ERR_ARG, "CSR column out of range");
    }
    s.col16 = c16;
    ESP_CHECK(values || h->grid_dev, ESPRESO_ERR_ARG, "null K values");
    // the values go into pinned memory too (uploaded from there at the set-up):
    // with sorted rows only the lower triangle, since K is symmetric (the
    // device mirrors the rest; half the bytes over the link)
    s.val_half = values && c16 && sorted;
    double *vp = values ? h->pin_alloc<double>(h->pin_val, s.val_half ? (size_t)nlow : (size_t)nnz, s.vpin_pos)
                        : nullptr;
    if (!vp) s.val_half = false;
    s.val_pin = vp;
    if (vp) {
        std::vector<double>().swap(s.vals);
        for (int q = 0; q < nnz; q++)
            ESP_CHECK(std::isfinite(values[q]), ESPRESO_ERR_NUMERIC, "non-finite value in K");
        if (s.val_half) s.lower_values(values, vp);
        else std::copy(values, values + nnz, vp);
    } else if (values) {
        s.vals.assign(values, values + nnz);
        for (int q = 0; q < nnz; q++)
            ESP_CHECK(std::isfinite(s.vals[q]), ESPRESO_ERR_NUMERIC, "non-finite value in K");
    } else {
        s.vals.clear();  // assembled on the device (structured grid)
    }
    s.f.assign(f, f + n);
    for (double v : s.f) ESP_CHECK(std::isfinite(v), ESPRESO_ERR_NUMERIC, "non-finite value in f");
    s.kdim = kdim;
    s.R.assign(R, R + (size_t)n * kdim);
    s.lam.assign(lambda, lambda + nlam);
    s.dof.assign(dof, dof + nlam);
    s.sign.assign(sign, sign + nlam);
    s.scale.assign(scale, scale + nlam);
    for (int q = 0; q < nlam; q++) {
        ESP_CHECK(s.dof[q] >= 0 && s.dof[q] < n, ESPRESO_ERR_ARG, "multiplier DOF out of range");
        ESP_CHECK(q == 0 || s.lam[q] > s.lam[q - 1], ESPRESO_ERR_ARG, "multipliers must be ascending");
    }
    s.set = true;
    ESP_API_END
}

int espreso_feti_set_cluster(espreso_feti *h, int32_t i, int32_t cluster, int32_t n0, const int32_t *row,
                             const int32_t *dof, const double *sign) {
    ESP_API_BEGIN
    ESP_CHECK(h && i >= 0 && i < h->nsub && cluster >= 0 && n0 >= 0, ESPRESO_ERR_ARG, "invalid cluster data");
    ESP_CHECK(!h->loaded, ESPRESO_ERR_STATE, "setup already done");
    ESP_CHECK(h->sub[i].set, ESPRESO_ERR_STATE, "set the subdomain before its cluster data");
    SubHost &s = h->sub[i];
    s.cluster = cluster;
    s.b0row.assign(row, row + n0);
    s.b0dof.assign(dof, dof + n0);
    s.b0sign.assign(sign, sign + n0);
    for (int q = 0; q < n0; q++) {
        ESP_CHECK(s.b0dof[q] >= 0 && s.b0dof[q] < s.n, ESPRESO_ERR_ARG, "B0 DOF out of range");
        ESP_CHECK(std::isfinite(s.b0sign[q]), ESPRESO_ERR_NUMERIC, "non-finite B0 value");
        ESP_CHECK(s.b0row[q] >= 0, ESPRESO_ERR_ARG, "B0 row must be non-negative");
    }
    ESP_API_END
}

int espreso_feti_set_dual(espreso_feti *h, int32_t n_lambda, const double *c) {
    ESP_API_BEGIN
    ESP_CHECK(h && n_lambda > 0 && c, ESPRESO_ERR_ARG, "invalid dual size");
    for (int32_t l = 0; l < n_lambda; l++) ESP_CHECK(std::isfinite(c[l]), ESPRESO_ERR_NUMERIC, "non-finite value in c");
    if (h->loaded) {
        // after the first set-up the multiplier layout is fixed: new Dirichlet
        // values of the same rows only (b is recomputed on the resident operators)
        ESP_CHECK(n_lambda == h->nlam, ESPRESO_ERR_STATE,
                  "espreso_feti_set_dual after setup: n_lambda differs from the loaded " + std::to_string(h->nlam));
        ESP_HIP(hipSetDevice(h->opt.device));
        h->c.assign(c, c + n_lambda);
        ESP_HIP(hipMemcpyAsync(h->d_c, h->c.data(), n_lambda * sizeof(double), hipMemcpyHostToDevice, h->st));
        if (h->ready) h->compute_b();
        h->sync();
        return ESPRESO_OK;
    }
    h->nlam = n_lambda;
    h->c.assign(c, c + n_lambda);
    ESP_API_END
}

int espreso_feti_update_values(espreso_feti *h, int32_t i, const double *values) {
    ESP_API_BEGIN
    ESP_CHECK(h && values && i >= 0 && i < h->nsub, ESPRESO_ERR_ARG, "invalid argument");
    SubHost &s = h->sub[i];
    ESP_CHECK(s.set, ESPRESO_ERR_STATE, "subdomain not set");
    const int nnz = s.rowptr[s.n];
    for (int q = 0; q < nnz; q++) ESP_CHECK(std::isfinite(values[q]), ESPRESO_ERR_NUMERIC, "non-finite value in K");
    if (!h->loaded) {
        if (s.val_pin && s.val_half) s.lower_values(values, s.val_pin);
        else if (s.val_pin) std::copy(values, values + nnz, s.val_pin);
        else s.vals.assign(values, values + nnz);
        return ESPRESO_OK;
    }
    ESP_HIP(hipSetDevice(h->opt.device));
    ESP_HIP(hipMemcpyAsync(h->P.vals + h->h_nzoff[i], values, nnz * sizeof(double), hipMemcpyHostToDevice, h->st));
    h->sync();
    if (s.val_pin && s.val_half) s.lower_values(values, s.val_pin);
    else if (s.val_pin) std::copy(values, values + nnz, s.val_pin);
    else if (!s.vals.empty()) s.vals.assign(values, values + nnz);
    h->values_dirty = true;
    h->ready = false;   // F_i, S_D and the coarse data follow at the next espreso_feti_setup
    h->solved = false;
    ESP_API_END
}

int espreso_feti_setup(espreso_feti *h) {
    ESP_API_BEGIN
    ESP_CHECK(h, ESPRESO_ERR_ARG, "null handle");
    ESP_HIP(hipSetDevice(h->opt.device));
    // a set-up that fails leaves no host thread reading the instance
    struct JoinGGt {
        espreso_feti *h;
        ~JoinGGt() {
            if (h->ggt_thread.joinable()) h->ggt_thread.join();
        }
    } join_ggt{h};
    if (!h->loaded) h->load();
    h->gpu_setup();
    ESP_API_END
}

int espreso_feti_set_rhs(espreso_feti *h, int32_t i, const double *f) {
    ESP_API_BEGIN
    ESP_CHECK(h && f && i >= 0 && i < h->nsub, ESPRESO_ERR_ARG, "invalid argument");
    ESP_CHECK(h->sub[i].set, ESPRESO_ERR_STATE, "subdomain not set");
    SubHost &s = h->sub[i];
    for (int q = 0; q < s.n; q++) ESP_CHECK(std::isfinite(f[q]), ESPRESO_ERR_ARG, "non-finite load");
    s.f.assign(f, f + s.n);
    ESP_API_END
}

int espreso_feti_resolve(espreso_feti *h, espreso_feti_stats *st) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready, ESPRESO_ERR_STATE, "espreso_feti_setup not called");
    ESP_HIP(hipSetDevice(h->opt.device));
    const double t0 = wall();
    h->resolve_rhs();
    const double trhs = wall() - t0;
    h->solve(st);
    if (st) st->time_coarse = trhs;
    if (!h->last.converged) return ESPRESO_NOT_CONVERGED;
    ESP_API_END
}

int espreso_feti_symv_items(espreso_feti *h, int64_t *targets) {
    ESP_API_BEGIN
    ESP_CHECK(h && targets, ESPRESO_ERR_ARG, "null argument");
    for (int p = 0; p < 3; p++) targets[p] = h->symv_target[p];
    ESP_API_END
}

int espreso_feti_subdomain_values(espreso_feti *h, int32_t i, double *vals) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready && vals, ESPRESO_ERR_STATE, "espreso_feti_setup not called");
    ESP_CHECK(i >= 0 && i < h->nsub, ESPRESO_ERR_ARG, "subdomain out of range");
    ESP_HIP(hipSetDevice(h->opt.device));
    const int64_t nnz = h->h_nzoff[i + 1] - h->h_nzoff[i];
    ESP_HIP(hipMemcpy(vals, h->P.vals + h->h_nzoff[i], nnz * sizeof(double), hipMemcpyDeviceToHost));
    ESP_API_END
}

int espreso_feti_resolve_block(espreso_feti *h, int32_t nrhs, const double *f, double *u, int32_t *iterations,
                               double *final_norms, espreso_feti_stats *stats) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready, ESPRESO_ERR_STATE, "espreso_feti_setup not called");
    ESP_CHECK(nrhs >= 0 && (nrhs == 0 || (f && u)), ESPRESO_ERR_ARG, "invalid block arguments");
    ESP_HIP(hipSetDevice(h->opt.device));
    const int64_t nf = (int64_t)nrhs * h->uoff[h->nsub];
    int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (int64_t q = 0; q < nf; q++) bad += !std::isfinite(f[q]);
    ESP_CHECK(bad == 0, ESPRESO_ERR_NUMERIC, "non-finite load");
    if (!h->resolve_block(nrhs, f, u, iterations, final_norms, stats)) return ESPRESO_NOT_CONVERGED;
    ESP_API_END
}

int espreso_feti_solve(espreso_feti *h, espreso_feti_stats *st) {
    ESP_API_BEGIN
    ESP_CHECK(h, ESPRESO_ERR_ARG, "null handle");
    ESP_HIP(hipSetDevice(h->opt.device));
    h->solve(st);
    if (!h->last.converged) return ESPRESO_NOT_CONVERGED;
    ESP_API_END
}

int espreso_feti_primal(espreso_feti *h, int32_t i, double *u) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->solved && i >= 0 && i < h->nsub, ESPRESO_ERR_STATE, "no solution");
    ESP_HIP(hipMemcpy(u, h->d_u + h->uoff[i], h->sub[i].n * sizeof(double), hipMemcpyDeviceToHost));
    ESP_API_END
}

int espreso_feti_load_work(espreso_feti *h, double *c) {
    ESP_API_BEGIN
    ESP_CHECK(h && c, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK(h->solved, ESPRESO_ERR_STATE, "no solution");
    ESP_HIP(hipSetDevice(h->opt.device));
    launch_dot(h->d_f, h->d_u, h->uoff[h->nsub], h->d_partial, h->d_scal + 8, h->st);
    ESP_HIP(hipMemcpyAsync(c, h->d_scal + 8, sizeof(double), hipMemcpyDeviceToHost, h->st));
    h->sync();
    ESP_API_END
}

int espreso_feti_dual(espreso_feti *h, double *lambda) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->solved, ESPRESO_ERR_STATE, "no solution");
    ESP_HIP(hipMemcpy(lambda, h->d_x, h->nlam * sizeof(double), hipMemcpyDeviceToHost));
    ESP_API_END
}

static int dual_op(espreso_feti *h, const double *x, double *y, int which) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready, ESPRESO_ERR_STATE, "setup not done");
    const size_t bytes = h->nlam * sizeof(double);
    ESP_HIP(hipMemcpyAsync(h->d_z, x, bytes, hipMemcpyHostToDevice, h->st));
    if (which == 0) h->apply_F(h->d_z, h->d_tmp, h->opt.sc_precision);
    else if (which == 1) h->apply_prec(h->d_z, h->d_tmp);
    else launch_project(h->D, h->d_z, h->d_tmp, 0, h->st);
    ESP_HIP(hipMemcpyAsync(y, h->d_tmp, bytes, hipMemcpyDeviceToHost, h->st));
    h->sync();
    ESP_API_END
}

int espreso_feti_apply_F(espreso_feti *h, const double *x, double *y) { return dual_op(h, x, y, 0); }
int espreso_feti_apply_prec(espreso_feti *h, const double *x, double *y) { return dual_op(h, x, y, 1); }
int espreso_feti_project(espreso_feti *h, const double *x, double *y) { return dual_op(h, x, y, 2); }

// ---- allocation-free projector ----
// Scratch of espreso_feti_project_device: the coarse vectors of one
// projection (G x per subdomain, G x per coarse block, (G G^T)^-1 G x).  Each
// GPU the library runs on holds its own copy for the life of the process;
// kernels get its device address (hipGetSymbolAddress, once per device),
// never the name.  Calls that overlap on one GPU share it and race.
constexpr int64_t PROJECT_SCRATCH_DOUBLES = (int64_t)1 << 20;  // 8 MiB
static __device__ __align__(256) double scratch[PROJECT_SCRATCH_DOUBLES];

static double *scratch_address(int device) {
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto f = addr.find(device);
    if (f != addr.end()) return f->second;
    void *p = nullptr;
    ESP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ESP_CHECK(p != nullptr, ESPRESO_ERR_HIP, "no device address for the projector scratch");
    addr[device] = (double *)p;
    return (double *)p;
}

int espreso_feti_project_device(espreso_feti *h, const double *x, double *y, void *stream) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready, ESPRESO_ERR_STATE, "setup not done");
    ESP_CHECK(x && y, ESPRESO_ERR_ARG, "null vector");
    ESP_HIP(hipSetDevice(h->opt.device));
    const int64_t need = h->n_dsub + 2 * (int64_t)h->ncoarse;
    ESP_CHECK(need <= PROJECT_SCRATCH_DOUBLES, ESPRESO_ERR_ARG,
              "coarse space too large for the allocation-free projector: it needs " + std::to_string(need) +
                  " doubles of scratch, the limit is " + std::to_string(PROJECT_SCRATCH_DOUBLES) +
                  " (use espreso_feti_project)");
    double *sp = scratch_address(h->opt.device);
    DualDev E = h->D;  // the instance's operators with the scratch in place of its coarse vectors
    E.dsub = sp;
    E.cd = sp + h->n_dsub;
    E.ce = sp + h->n_dsub + h->ncoarse;
    launch_project(E, x, y, 0, stream ? (hipStream_t)stream : h->st);
    h->check_launches("projector");
    ESP_API_END
}

int espreso_feti_vec_b(espreso_feti *h, double *b) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready, ESPRESO_ERR_STATE, "setup not done");
    ESP_HIP(hipMemcpy(b, h->d_b, h->nlam * sizeof(double), hipMemcpyDeviceToHost));
    ESP_API_END
}

int espreso_feti_subdomain_boundary(espreso_feti *h, int32_t i, int32_t *m, int32_t *bdofs) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready && i >= 0 && i < h->nsub, ESPRESO_ERR_STATE, "setup not done");
    *m = h->sub[i].m;
    if (bdofs) std::copy(h->sub[i].bdofs.begin(), h->sub[i].bdofs.end(), bdofs);
    ESP_API_END
}

int espreso_feti_subdomain_dense(espreso_feti *h, int32_t i, int32_t which, double *out) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready && i >= 0 && i < h->nsub, ESPRESO_ERR_STATE, "setup not done");
    const int m = h->sub[i].m;
    double *tmp = h->alloc<double>((size_t)m * m);
    std::vector<int64_t> off(2);
    ESP_HIP(hipMemcpy(off.data(), h->D.foff + i, 2 * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (which == 0) launch_unpack(h->D.Fd, 0, off[0], m, 1.0f, tmp, h->st);
    else if (which == 1) {
        const int sp = h->D.sd_prec;
        std::vector<int64_t> lo(1, 0);
        float sc = 1.0f;
        if (sp) {
            ESP_HIP(hipMemcpy(lo.data(), h->D.poff[sp] + i, sizeof(int64_t), hipMemcpyDeviceToHost));
            ESP_HIP(hipMemcpy(&sc, h->D.sdscale + i, sizeof(float), hipMemcpyDeviceToHost));
        }
        if (sp == 0) launch_unpack(h->D.SD, 0, off[0], m, 1.0f, tmp, h->st);
        else if (sp == 1) launch_unpack(h->D.SDf, 1, lo[0], m, 1.0f, tmp, h->st);
        else launch_unpack(h->D.SDh, 2, lo[0], m, sc, tmp, h->st);
    } else {
        const int prec = h->opt.sc_precision;
        std::vector<int64_t> lo(1);
        float sc = 1.0f;
        if (prec) {
            ESP_HIP(hipMemcpy(lo.data(), h->D.poff[prec] + i, sizeof(int64_t), hipMemcpyDeviceToHost));
            ESP_HIP(hipMemcpy(&sc, h->D.fscale + i, sizeof(float), hipMemcpyDeviceToHost));
        }
        if (prec == 0) launch_unpack(h->D.Fd, 0, off[0], m, 1.0f, tmp, h->st);
        else if (prec == 1) launch_unpack(h->D.Ff, 1, lo[0], m, 1.0f, tmp, h->st);
        else launch_unpack(h->D.Fh, 2, lo[0], m, sc, tmp, h->st);
    }
    ESP_HIP(hipMemcpyAsync(out, tmp, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost, h->st));
    h->sync();
    h->release(tmp);
    ESP_API_END
}

int espreso_feti_time_apply(espreso_feti *h, int32_t op, int32_t reps, double *ms_apply, double *ms_symv) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->ready && reps > 0, ESPRESO_ERR_STATE, "setup not done");
    hipEvent_t e0, e1, s0, s1;
    ESP_HIP(hipEventCreate(&e0));
    ESP_HIP(hipEventCreate(&e1));
    ESP_HIP(hipEventCreate(&s0));
    ESP_HIP(hipEventCreate(&s1));
    launch_axpby(h->d_z, h->d_b, h->d_b, h->nlam, 1.0, 0.0, h->st);
    // warm-up
    if (op == 0) h->apply_F(h->d_z, h->d_tmp, h->opt.sc_precision);
    else h->apply_prec(h->d_z, h->d_tmp);
    double symv_ms = 0;
    ESP_HIP(hipEventRecord(e0, h->st));
    for (int r = 0; r < reps; r++) {
        if (op == 0) h->apply_F(h->d_z, h->d_tmp, h->opt.sc_precision);
        else h->apply_prec(h->d_z, h->d_tmp);
    }
    ESP_HIP(hipEventRecord(e1, h->st));
    ESP_HIP(hipEventSynchronize(e1));
    float tot = 0;
    ESP_HIP(hipEventElapsedTime(&tot, e0, e1));
    for (int r = 0; r < reps; r++) {
        launch_symv(h->D, op != 0, op == 0 ? h->opt.sc_precision : 0, h->d_z, 1, op != 0, h->st, s0, s1);
        ESP_HIP(hipEventSynchronize(s1));
        float t = 0;
        ESP_HIP(hipEventElapsedTime(&t, s0, s1));
        symv_ms += t;
    }
    *ms_apply = tot / reps;
    *ms_symv = symv_ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(s0);
    (void)hipEventDestroy(s1);
    ESP_API_END
}

int espreso_debug_copy(int64_t bytes, int32_t reps, double *ms) {
    ESP_API_BEGIN
    ESP_CHECK(bytes >= 32 && reps > 0, ESPRESO_ERR_ARG, "copy size must be >= 32 bytes");
    const int64_t n = bytes / 32 * 4;
    double *a = nullptr, *b = nullptr;
    ESP_HIP(hipMalloc(&a, n * sizeof(double)));
    hipError_t e = hipMalloc(&b, n * sizeof(double));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipMemset(a, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float t = 0;
    if (e == hipSuccess) {
        launch_copy(a, b, n, nullptr);
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; r++) launch_copy(a, b, n, nullptr);
        (void)hipEventRecord(e1, nullptr);
        e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    ESP_HIP(e);
    *ms = t / reps;
    ESP_API_END
}

int espreso_debug_stream(int64_t bytes, int32_t reps, int32_t workgroups, double *ms) {
    ESP_API_BEGIN
    ESP_CHECK(bytes >= 16 && reps > 0 && workgroups > 0 && workgroups <= 1 << 20, ESPRESO_ERR_ARG,
              "invalid stream arguments");
    const int64_t n = bytes / 16 * 2;
    double *a = nullptr, *o = nullptr;
    ESP_HIP(hipMalloc(&a, n * sizeof(double)));
    hipError_t e = hipMalloc(&o, (size_t)workgroups * sizeof(double));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipMemset(a, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float t = 0;
    if (e == hipSuccess) {
        launch_stream_read(a, n, workgroups, o, nullptr);
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; r++) launch_stream_read(a, n, workgroups, o, nullptr);
        (void)hipEventRecord(e1, nullptr);
        e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (o) (void)hipFree(o);
    ESP_HIP(e);
    *ms = t / reps;
    ESP_API_END
}

int espreso_debug_mfma_rate(int32_t shape, int32_t chains, int32_t workgroups_per_cu, double *tflops) {
    ESP_API_BEGIN
    const bool chains_ok = shape <= 1   ? (chains == 4 || chains == 8 || chains == 16 || chains == 32)
                           : shape == 2 ? (chains == 2 || chains == 4 || chains == 8)
                                        : (chains == 2 || chains == 4);
    ESP_CHECK(shape >= 0 && shape <= 3 && chains_ok && tflops && workgroups_per_cu >= 0 && workgroups_per_cu <= 8,
              ESPRESO_ERR_ARG,
              "shape 0..3; chains 4, 8, 16 or 32 (shapes 0, 1), 2, 4 or 8 (shape 2), 2 or 4 (shape 3); "
              "0..8 workgroups per CU");
    const int blocks = 4096, iter = 2000;
    // unused dynamic LDS that lets only workgroups_per_cu workgroups share a CU
    int lds = 0;
    if (workgroups_per_cu > 0) {
        int cu_lds = 0;
        ESP_HIP(hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, 0));
        lds = cu_lds / workgroups_per_cu - 1024;
        int blk_lds = 0;
        ESP_HIP(hipDeviceGetAttribute(&blk_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, 0));
        ESP_CHECK(lds > 0 && lds <= blk_lds, ESPRESO_ERR_ARG, "workgroups per CU not reachable by the LDS size");
    }
    double *o = nullptr;
    ESP_HIP(hipMalloc(&o, (size_t)blocks * 256 * sizeof(double)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float t = 0;
    if (e == hipSuccess) {
        launch_mfma_rate(shape, chains, blocks, 16, lds, o, nullptr);  // warm
        (void)hipEventRecord(e0, nullptr);
        launch_mfma_rate(shape, chains, blocks, iter, lds, o, nullptr);
        (void)hipEventRecord(e1, nullptr);
        e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(o);
    ESP_HIP(e);
    ESP_CHECK(t > 0, ESPRESO_ERR_STATE, "MFMA rate loop not timed");
    const double per_iter = shape >= 2 ? (double)chains * chains : (double)chains;  // MFMAs per wave and step
    *tflops = (double)blocks * 4 * iter * per_iter * mfma_rate_flops(shape) / (t * 1e-3) / 1e12;
    ESP_API_END
}

// ---- diagnostics of the MFMA tile routine ----
int espreso_debug_mfma(const double *a, const double *b, double *out) {
    ESP_API_BEGIN
    double *d = nullptr;
    ESP_HIP(hipMalloc(&d, (64 + 64 + 320) * sizeof(double)));
    hipError_t e = hipMemcpy(d, a, 64 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 64, b, 64 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_probe_mfma(d, d + 64, d + 128, nullptr);
        e = hipMemcpy(out, d + 128, 320 * sizeof(double), hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    ESP_HIP(e);
    ESP_API_END
}

int espreso_debug_tile_gemm(int32_t ta, int32_t tb, int32_t M, int32_t N, int32_t K, const double *A, int32_t lda,
                            const double *B, int32_t ldb, double *C, int32_t ldc, double alpha, double beta) {
    ESP_API_BEGIN
    ESP_CHECK(M >= 0 && M <= 64 && N >= 0 && N <= 64 && K >= 0, ESPRESO_ERR_ARG, "tile GEMM shape out of range");
    const size_t na = (size_t)lda * (ta ? M : K), nbb = (size_t)ldb * (tb ? K : N), nc = (size_t)ldc * N;
    ESP_CHECK(lda >= (ta ? K : M) && ldb >= (tb ? N : K) && ldc >= M, ESPRESO_ERR_ARG, "leading dimension too small");
    double *d = nullptr;
    ESP_HIP(hipMalloc(&d, (na + nbb + nc + 1) * sizeof(double)));
    hipError_t e = hipMemcpy(d, A, na * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + na, B, nbb * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + na + nbb, C, nc * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_probe_gemm(ta, tb, M, N, K, d, lda, d + na, ldb, d + na + nbb, ldc, alpha, beta, nullptr);
        e = hipMemcpy(C, d + na + nbb, nc * sizeof(double), hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    ESP_HIP(e);
    ESP_API_END
}

// ---- structured grid front end ----

int espreso_grid_sizes(const espreso_grid_problem *p, int32_t *n_subdomains, int32_t *n_nodes, int32_t *ndof) {
    ESP_API_BEGIN
    ESP_CHECK(p, ESPRESO_ERR_ARG, "null problem");
    *n_subdomains = grid_subdomain_count(*p);
    *n_nodes = grid_node_count(*p);
    *ndof = p->physics == 0 ? 3 : 1;
    ESP_API_END
}

int espreso_grid_load(espreso_feti *h, const espreso_grid_problem *p) {
    ESP_API_BEGIN
    ESP_CHECK(h && p, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK(grid_subdomain_count(*p) == h->nsub, ESPRESO_ERR_ARG, "subdomain count differs from the instance");
    GridProblem g;
    h->grid_dev = p->device_assembly != 0;
    const double tb = wall();
    grid_build(*p, h->opt.scaling, g, !h->grid_dev);
    h->last.time_build = wall() - tb;
    h->ndof = g.ndof;
    h->nnodes = g.nnodes;
    for (int d = 0; d < 3; d++) h->gridNE[d] = g.NN[d] - 1;
    h->gridKe0 = g.Ke;
    h->grid_box.resize(6 * (size_t)h->nsub);
    for (int s = 0; s < h->nsub; s++) std::copy(g.sub[s].box, g.sub[s].box + 6, h->grid_box.begin() + 6 * s);
    h->grid_escale.swap(g.escale);
    h->l2g.resize(h->nsub);
    // the subdomains are handed over through the public entry points, in
    // parallel (each call touches its own subdomain only)
    std::vector<int> rcs(h->nsub, 0);
    std::vector<std::string> errs(h->nsub);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < h->nsub; s++) {
        GridSubdomain &sd = g.sub[s];
        int rc = espreso_feti_set_subdomain(h, s, sd.n, sd.rowptr.data(), sd.colidx.data(),
                                            h->grid_dev ? nullptr : sd.vals.data(),
                                            sd.f.data(), sd.kdim, sd.R.data(), (int)sd.lam.size(), sd.lam.data(),
                                            sd.dof.data(), sd.sign.data(), sd.scale.data());
        if (!rc && h->opt.method == 1)
            rc = espreso_feti_set_cluster(h, s, sd.cluster, (int)sd.b0row.size(), sd.b0row.data(), sd.b0dof.data(),
                                          sd.b0sign.data());
        if (rc) {
            rcs[s] = rc;
            errs[s] = espreso_last_error();
            continue;
        }
        h->l2g[s].swap(sd.l2g);
        GridSubdomain().rowptr.swap(sd.rowptr);
        std::vector<int>().swap(sd.colidx);
        std::vector<double>().swap(sd.vals);
    }
    for (int s = 0; s < h->nsub; s++)
        if (rcs[s]) throw Error(rcs[s], errs[s]);
    return espreso_feti_set_dual(h, g.nlam, g.c.data());
    ESP_API_END
}

int espreso_grid_export(const espreso_grid_problem *p, int32_t scaling, int32_t i, int32_t *sizes, int32_t *rowptr,
                        int32_t *colidx, double *values, double *f, double *R, int32_t *lambda, int32_t *dof,
                        double *sign, double *scale, double *c) {
    ESP_API_BEGIN
    ESP_CHECK(p && sizes, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK(i >= 0 && i < grid_subdomain_count(*p), ESPRESO_ERR_ARG, "subdomain out of range");
    GridProblem g;
    grid_build(*p, scaling, g, p->device_assembly == 0);
    const GridSubdomain &sd = g.sub[i];
    sizes[0] = sd.n;
    sizes[1] = sd.rowptr[sd.n];
    sizes[2] = sd.kdim;
    sizes[3] = (int)sd.lam.size();
    sizes[4] = g.nlam;
    sizes[5] = (int)g.sub.size();
    if (rowptr) std::copy(sd.rowptr.begin(), sd.rowptr.end(), rowptr);
    if (colidx) std::copy(sd.colidx.begin(), sd.colidx.end(), colidx);
    if (values) {
        if (sd.vals.empty()) std::fill(values, values + sd.rowptr[sd.n], 0.0);  // device-assembled values
        else std::copy(sd.vals.begin(), sd.vals.end(), values);
    }
    if (f) std::copy(sd.f.begin(), sd.f.end(), f);
    if (R) std::copy(sd.R.begin(), sd.R.end(), R);
    if (lambda) std::copy(sd.lam.begin(), sd.lam.end(), lambda);
    if (dof) std::copy(sd.dof.begin(), sd.dof.end(), dof);
    if (sign) std::copy(sd.sign.begin(), sd.sign.end(), sign);
    if (scale) std::copy(sd.scale.begin(), sd.scale.end(), scale);
    if (c) std::copy(g.c.begin(), g.c.end(), c);
    ESP_API_END
}

int espreso_grid_export_cluster(const espreso_grid_problem *p, int32_t i, int32_t *cluster, int32_t *n0,
                                int32_t *row, int32_t *dof, double *sign) {
    ESP_API_BEGIN
    ESP_CHECK(p && cluster && n0, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK(i >= 0 && i < grid_subdomain_count(*p), ESPRESO_ERR_ARG, "subdomain out of range");
    GridProblem g;
    grid_build(*p, 1, g);
    const GridSubdomain &sd = g.sub[i];
    *cluster = sd.cluster;
    *n0 = (int)sd.b0row.size();
    if (row) std::copy(sd.b0row.begin(), sd.b0row.end(), row);
    if (dof) std::copy(sd.b0dof.begin(), sd.b0dof.end(), dof);
    if (sign) std::copy(sd.b0sign.begin(), sd.b0sign.end(), sign);
    ESP_API_END
}

int espreso_grid_nodal_load(espreso_feti *h, double *g) {
    ESP_API_BEGIN
    ESP_CHECK(h && g, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK((int)h->l2g.size() == h->nsub, ESPRESO_ERR_STATE, "grid not loaded");
    const int nd = h->ndof;
    std::fill(g, g + (size_t)h->nnodes * nd, 0.0);
    for (int s = 0; s < h->nsub; s++)
        for (size_t ln = 0; ln < h->l2g[s].size(); ln++)
            for (int c = 0; c < nd; c++) g[(size_t)h->l2g[s][ln] * nd + c] += h->sub[s].f[ln * nd + c];
    ESP_API_END
}

int espreso_grid_unit_element(const espreso_grid_problem *p, double *Ke) {
    ESP_API_BEGIN
    ESP_CHECK(p && Ke, ESPRESO_ERR_ARG, "null argument");
    std::vector<double> K;
    grid_unit_element(*p, K);
    std::copy(K.begin(), K.end(), Ke);
    ESP_API_END
}

int espreso_grid_set_nodal_rhs(espreso_feti *h, const double *g) {
    ESP_API_BEGIN
    ESP_CHECK(h && g, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK((int)h->l2g.size() == h->nsub, ESPRESO_ERR_STATE, "grid not loaded");
    const int nd = h->ndof;
    std::vector<int> mult(h->nnodes, 0);
    for (int s = 0; s < h->nsub; s++)
        for (int node : h->l2g[s]) mult[node]++;
    for (int s = 0; s < h->nsub; s++) {
        std::vector<double> f(h->sub[s].n);
        for (size_t ln = 0; ln < h->l2g[s].size(); ln++) {
            const int node = h->l2g[s][ln];
            for (int c = 0; c < nd; c++) f[ln * nd + c] = g[(size_t)node * nd + c] / mult[node];
        }
        const int rc = espreso_feti_set_rhs(h, s, f.data());
        if (rc) return rc;
    }
    ESP_API_END
}

int espreso_grid_resolve_block(espreso_feti *h, int32_t nrhs, const double *g, double *u_nodal, int32_t *iterations,
                               double *final_norms, espreso_feti_stats *stats) {
    ESP_API_BEGIN
    ESP_CHECK(h && (nrhs == 0 || (g && u_nodal)) && nrhs >= 0, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK((int)h->l2g.size() == h->nsub, ESPRESO_ERR_STATE, "grid not loaded");
    ESP_CHECK(h->ready, ESPRESO_ERR_STATE, "espreso_feti_setup not called");
    const int nd = h->ndof, nsub = h->nsub;
    const int64_t nt = h->uoff[nsub], nv = (int64_t)h->nnodes * nd;
    std::vector<int> mult(h->nnodes, 0);
    for (int s = 0; s < nsub; s++)
        for (int node : h->l2g[s]) mult[node]++;
    // loads split over the subdomain copies of a node (as espreso_grid_set_nodal_rhs)
    std::vector<double> f((size_t)nrhs * nt), u((size_t)nrhs * nt);
#pragma omp parallel for collapse(2) schedule(dynamic, 8)
    for (int v = 0; v < nrhs; v++)
        for (int s = 0; s < nsub; s++) {
            double *fs = f.data() + v * nt + h->uoff[s];
            const double *gv = g + v * nv;
            for (size_t ln = 0; ln < h->l2g[s].size(); ln++) {
                const int node = h->l2g[s][ln];
                for (int c = 0; c < nd; c++) fs[ln * nd + c] = gv[(size_t)node * nd + c] / mult[node];
            }
        }
    int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (int64_t q = 0; q < (int64_t)f.size(); q++) bad += !std::isfinite(f[q]);
    ESP_CHECK(bad == 0, ESPRESO_ERR_NUMERIC, "non-finite load");
    const bool conv = h->resolve_block(nrhs, f.data(), u.data(), iterations, final_norms, stats);
    // solutions averaged over the subdomain copies (as espreso_grid_nodal_solution)
#pragma omp parallel for schedule(static)
    for (int v = 0; v < nrhs; v++) {
        double *un = u_nodal + v * nv;
        std::fill(un, un + nv, 0.0);
        for (int s = 0; s < nsub; s++) {
            const double *us = u.data() + v * nt + h->uoff[s];
            for (size_t ln = 0; ln < h->l2g[s].size(); ln++) {
                const int node = h->l2g[s][ln];
                for (int c = 0; c < nd; c++) un[(size_t)node * nd + c] += us[ln * nd + c];
            }
        }
        for (int node = 0; node < h->nnodes; node++)
            for (int c = 0; c < nd; c++) un[(size_t)node * nd + c] /= mult[node];
    }
    if (!conv) return ESPRESO_NOT_CONVERGED;
    ESP_API_END
}

int espreso_grid_update_material(espreso_feti *h, const double *material_elem) {
    ESP_API_BEGIN
    ESP_CHECK(h && material_elem, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK(h->gridNE[0] > 0 && !h->gridKe0.empty(), ESPRESO_ERR_STATE, "grid not loaded");
    const int64_t nel = (int64_t)h->gridNE[0] * h->gridNE[1] * h->gridNE[2];
    ESP_CHECK((int64_t)h->grid_escale.size() == nel, ESPRESO_ERR_STATE, "grid materials missing");
    for (int64_t e = 0; e < nel; e++)
        ESP_CHECK(std::isfinite(material_elem[e]) && material_elem[e] > 0, ESPRESO_ERR_ARG,
                  "element material must be positive and finite");
    std::copy(material_elem, material_elem + nel, h->grid_escale.begin());
    if (!h->loaded) {
        // the values are computed on the device at the first set-up
        h->grid_dev = true;
        for (SubHost &s : h->sub) {
            std::vector<double>().swap(s.vals);
            s.val_pin = nullptr;
        }
        return ESPRESO_OK;
    }
    ESP_HIP(hipSetDevice(h->opt.device));
    h->grid_stiffness();
    for (SubHost &s : h->sub) {
            std::vector<double>().swap(s.vals);
            s.val_pin = nullptr;
        }
    h->values_dirty = true;
    h->ready = false;
    h->solved = false;
    ESP_API_END
}

int espreso_grid_element_sensitivity(espreso_feti *h, const double *u, const double *w, double *sens) {
    ESP_API_BEGIN
    ESP_CHECK(h && u && w && sens, ESPRESO_ERR_ARG, "null argument");
    ESP_CHECK(h->gridNE[0] > 0 && !h->gridKe0.empty(), ESPRESO_ERR_STATE, "grid not loaded");
    ESP_HIP(hipSetDevice(h->opt.device));
    const int64_t nel = (int64_t)h->gridNE[0] * h->gridNE[1] * h->gridNE[2];
    const int64_t nv = (int64_t)h->nnodes * h->ndof;
    double *du = h->alloc<double>(nv), *dw = h->alloc<double>(nv), *ds = h->alloc<double>(nel);
    double *dk = h->upload(h->gridKe0);
    ESP_HIP(hipMemcpyAsync(du, u, nv * sizeof(double), hipMemcpyHostToDevice, h->st));
    ESP_HIP(hipMemcpyAsync(dw, w, nv * sizeof(double), hipMemcpyHostToDevice, h->st));
    launch_element_sensitivity(h->gridNE, h->ndof, dk, du, dw, ds, h->st);
    ESP_HIP(hipMemcpyAsync(sens, ds, nel * sizeof(double), hipMemcpyDeviceToHost, h->st));
    h->sync();
    for (void *q : {(void *)du, (void *)dw, (void *)ds, (void *)dk}) h->release(q);
    ESP_API_END
}

int espreso_grid_l2g(espreso_feti *h, int32_t i, int32_t *l2g) {
    ESP_API_BEGIN
    ESP_CHECK(h && i >= 0 && i < (int)h->l2g.size(), ESPRESO_ERR_STATE, "grid not loaded");
    std::copy(h->l2g[i].begin(), h->l2g[i].end(), l2g);
    ESP_API_END
}

int espreso_grid_nodal_solution(espreso_feti *h, double *u_nodal) {
    ESP_API_BEGIN
    ESP_CHECK(h && h->solved && !h->l2g.empty(), ESPRESO_ERR_STATE, "no grid solution");
    const int nd = h->ndof;
    std::vector<double> u(h->uoff[h->nsub]);
    ESP_HIP(hipMemcpy(u.data(), h->d_u, u.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> cnt(h->nnodes, 0);
    std::fill(u_nodal, u_nodal + (size_t)h->nnodes * nd, 0.0);
    for (int s = 0; s < h->nsub; s++)
        for (size_t ln = 0; ln < h->l2g[s].size(); ln++) {
            const int g = h->l2g[s][ln];
            cnt[g]++;
            for (int d = 0; d < nd; d++) u_nodal[(size_t)g * nd + d] += u[h->uoff[s] + ln * nd + d];
        }
    for (int g = 0; g < h->nnodes; g++)
        for (int d = 0; d < nd; d++) u_nodal[(size_t)g * nd + d] /= std::max(cnt[g], 1);
    ESP_API_END
}

}  // extern "C"
