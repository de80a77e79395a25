// Serving engine: continuous batching of image-formation jobs on one GPU.
//
// Jobs arrive on host threads (sarbp_server_submit) and are queued in arrival order. A scheduler
// thread admits them while their device memory fits the HBM budget and the in-flight limit allows;
// a job that does not fit lets smaller jobs behind it pass, until it has waited kHeadPatienceNs,
// after which nothing passes it. The budget is one device allocation made when the server is
// created and divided among admitted jobs by the scheduler (first fit, coalescing), so admission
// never calls the driver's allocator and jobs on different streams never wait on each other's frees.
//
// Work flows through streams with one role each, so uploads, kernels and downloads of different
// jobs overlap:
//   stage    uploads of every job (copies only): a small job's inputs at admission, a larger
//            job's phase history and pixel locations in chunks, at most kUploadChunksAhead
//            outstanding, so that a small job's upload waits for at most that many chunks
//   prep     pulse table and range profiles of small jobs (nu * nv * npulses at most the fusion
//            size)
//   big[i]   pulse table, range profiles and accumulation of larger jobs, in row bands (the least
//            loaded of num_streams streams)
//   fuse     accumulation of small jobs: prepared small jobs gather while every launch-table slot
//            is in flight, and when one is free, they are computed by one launch over a
//            job-indexed tile list (up to the table's capacity and fuse_launch_pixel_pulses)
//   dl_small, dl_big   output downloads to the caller's buffers
// A step that depends on another stream's work is enqueued by the scheduler once it has seen that
// work complete (event query), never as a device-side wait: streams share the process's few
// hardware queues, which run their packets in order, and a packet waiting there on work enqueued
// later would hold every stream mapped to the same queue. A job's arena space returns to the free
// list when the scheduler sees its download complete.
//
// A per-shape-class autotuner (shape class = N_fft, range-profile format, log2 of pulses and of
// pixels) times the tile candidates on the first job of a class, over a band of its rows, each twice
// in mirrored order, and uses the fastest for later jobs of that class; until the timings are in,
// jobs use the library's size-based default tile. A candidate is a tile height (16 x 32, 48 or 64)
// and, for 16 x 48 and 16 x 64, the pulse-block form of the accumulation: row elements gathered
// from HBM in blocks of 32 pulses, or row windows staged in LDS in blocks of 8 (the same bits).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "internal.h"

using namespace sarbp;

namespace {

constexpr int64_t kHeadPatienceNs = 50'000'000;     // 50 ms
constexpr size_t kFuseSlotBytes = (size_t)4 << 20;  // launch table (job args + tiles) per slot
constexpr int kFuseSlots = 2;                       // fused launches in flight at most
// Pixels per accumulation launch of a large job. While one launch of some 10^5 tiles runs, the
// device starts no kernel from any other queue (measured with a kernel trace), so a large job is
// computed in row bands and other jobs' kernels and copies are dispatched between bands. A band
// holds this many pixels whatever the tile height: 4096 tiles of 16 x 32 (two device waves at 8
// tiles per CU) or 2048 of 16 x 64 (two waves at 4 per CU), so that its duration does not grow with
// the tile the tuner picks.
constexpr int64_t kBandPixels = 4096 * 16 * 32;
constexpr int kBandsAhead = 2;
// A banded job that has enqueued no band for this long goes ahead of the least-work-left order
// (oldest first), so that no job waits without bound behind a run of smaller ones.
constexpr int64_t kBandPatienceNs = 50'000'000;  // 50 ms
// Four chunks in flight (~0.6 ms of copying) keep the copy engine busy across a scheduler pass
// over many jobs; with two, a large job's output took 2-5x as long when 2000 jobs were in flight.
constexpr size_t kUploadChunkBytes = (size_t)8 << 20;  // ~0.15 ms over PCIe 5 x16
constexpr int kUploadChunksAhead = 4;
// Pixel-pulses per fused launch at most by default (~4 ms of accumulation on an idle device), so
// that a fused launch holds the device no longer than a band does (sarbp_server_config.
// fuse_launch_pixel_pulses).
constexpr int64_t kFusePixelPulses = 8'000'000'000LL;
// Blocks of the range-profile launches of the engine (256 threads each, looping over pulses).
constexpr int kSharedFftBlocks = 512;

int64_t now_ns(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() -
                                                                t0)
        .count();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int ilog2(int64_t v) {
    int l = 0;
    while (v > 1) {
        v >>= 1;
        ++l;
    }
    return l;
}

// First-fit allocator over [0, size) with coalescing of free neighbours; offsets are 256-aligned.
class DeviceArena {
  public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    size_t alloc(size_t bytes) {
        bytes = align256(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = align256(bytes);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }

  private:
    std::map<size_t, size_t> free_;  // offset -> length
};

struct Layout {  // byte offsets of a job's device buffers inside its arena block
    size_t phs, k_r, pos, u, v, ws, wp, r, q, qscale, rot, pc, hdr, out, total;
    size_t host_lo, host_hi;  // [k_r, wp end): the block staged through page-locked host memory
};

// Page-locked host blocks in power-of-two size classes, reused across jobs: copies from them are
// asynchronous DMA, where a copy from pageable memory would hold the scheduler until it is staged.
class PinnedPool {
  public:
    void *get(size_t bytes, size_t *cls_out) {
        size_t cls = 4096;
        while (cls < bytes) cls <<= 1;
        *cls_out = cls;
        auto &lst = free_[cls];
        if (!lst.empty()) {
            void *p = lst.back();
            lst.pop_back();
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, cls, hipHostMallocDefault) != hipSuccess) return nullptr;
        return p;
    }
    void put(void *p, size_t cls) {
        if (p) free_[cls].push_back(p);
    }
    ~PinnedPool() {
        for (auto &kv : free_)
            for (void *p : kv.second) (void)hipHostFree(p);
    }

  private:
    std::map<size_t, std::vector<void *>> free_;
};

// Where an admitted job is: each step ends with an event the scheduler queries before the next.
enum class Phase {
    Uploading,    // copies on the stage stream, then ev_up
    Prepping,     // small: pulse table and range profiles on the prep stream, then ev_ready
    Ready,        // small: prepared, waiting in fuse_pending
    Computing,    // accumulation in a fused launch (slot event) or in bands on a big stream
    Finishing,    // large: all bands enqueued, then ev_acc
    Downloading,  // output copy on a download stream, then ev_done
};

struct Job {
    uint64_t id = 0;
    sarbp_job d{};
    RowLayout lay{};
    Layout ar{};
    std::vector<double> ws, wp;  // host windows (computed at submission)
    void *stage = nullptr;       // page-locked copy of k_r, pos, u, v, ws, wp
    size_t stage_cls = 0;
    int64_t t_submit = 0, t_admit = 0, t_end = 0;
    // when the scheduler saw each step end or began it (0: not reached): inputs uploaded, range
    // profiles computed, accumulation enqueued (fused launch or first band), accumulation complete
    int64_t t_up = 0, t_ready = 0, t_launch = 0, t_acc = 0;
    int64_t pp = 0;  // pixel-pulses of the job
    int stream = 0;
    int ppt = 0;
    int row_windows = 0;  // bands stage row windows in LDS (the tuner's choice)
    bool fused = false;
    bool small = false;
    bool prepared = false;       // image plane, pulse table and range profiles enqueued
    bool done_recorded = false;  // ev_done is recorded (a never-recorded event queries as complete)
    int status = 1;              // 1 queued, 2 admitted, 0 done, < 0 error
    Phase phase = Phase::Uploading;
    size_t off = SIZE_MAX;       // arena block
    uint8_t *dev = nullptr;
    hipEvent_t ev_up = nullptr;     // inputs in device memory
    bool up_recorded = false;       // ev_up is recorded (after a chunked upload's last chunk)
    size_t up_sent = 0, up_total = 0;            // larger jobs: bytes of phs and r enqueued, of all
    size_t dl_sent = 0;                          // larger jobs: bytes of the output enqueued
    hipEvent_t ev_chunk[kUploadChunksAhead] = {};  // larger jobs: the outstanding copy chunks
    int chunks_enqueued = 0, chunks_done = 0;
    hipEvent_t ev_ready = nullptr;  // small: prepared for the fused launch
    hipEvent_t ev_acc = nullptr;    // large: accumulated
    hipEvent_t ev_band[2] = {};     // large: completion of the outstanding accumulation bands
    bool banding = false;           // large: bands still to enqueue
    int next_row = 0, bands_enqueued = 0, bands_done = 0;
    int64_t t_band = 0;  // when the job last enqueued a band (or began banding)
    hipEvent_t ev_done = nullptr;   // output in host memory
    hipEvent_t ev_begin = nullptr;  // timing pair with ev_done: device time of the job
    int64_t host_ns = 0;            // scheduler time spent enqueuing the job's work
    float device_ms = 0.f;
    BpArgs args{};
    std::tuple<int, int, int, int> cls{};
};

struct FuseSlot {
    void *host = nullptr;  // page-locked launch table
    void *dev = nullptr;   // its device copy
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Job *> group;  // jobs of the launch in flight
};

// A tuning candidate: tile_ppt plus kTuneWindows when the job's bands stage row windows in LDS.
constexpr int kTuneWindows = 16;

struct Tuning {
    int best = 0;          // chosen candidate (tile_ppt | kTuneWindows; 0: not decided)
    bool running = false;  // candidate timings in flight
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> probes;
};

}  // namespace

struct sarbp_server {
    sarbp_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::vector<std::shared_ptr<Job>> inflight;
    std::map<uint64_t, std::shared_ptr<Job>> jobs;  // all jobs not yet collected by wait
    std::vector<Job *> fuse_pending;                // small jobs prepared, waiting for a fused launch
    FuseSlot slots[kFuseSlots];
    PinnedPool pinned;
    DeviceArena arena;
    uint8_t *arena_base = nullptr;
    std::map<std::tuple<int, int, int, int>, Tuning> tuning;
    std::vector<hipStream_t> streams;  // big[i]
    hipStream_t stage = nullptr, prep = nullptr, fuse_stream = nullptr, dl_small = nullptr,
                dl_big = nullptr;
    std::vector<int64_t> stream_load;  // outstanding pixel-pulses per big stream
    uint64_t next_id = 1;
    int64_t used_bytes = 0;
    bool stop = false;
    std::thread worker;
    const float2 *tw[16] = {};  // twiddle tables per log2 N_fft
};

namespace {

Layout layout_of(const sarbp_job &j, const RowLayout &lay) {
    Layout a{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off += align256(bytes);
        return at;
    };
    const size_t np = (size_t)j.npulses, ns = (size_t)j.nsamples;
    const size_t npix = (size_t)j.nu * (size_t)j.nv;
    a.phs = take(8 * np * ns);
    a.k_r = take(4 * ns);
    a.pos = take(12 * np);
    a.u = take(8 * (size_t)j.nu);
    a.v = take(8 * (size_t)j.nv);
    a.ws = take(8 * ns);
    a.wp = take(8 * np);
    a.host_lo = a.k_r;
    a.host_hi = a.wp + 8 * np;
    a.r = take(j.pixel_locs ? 24 * npix : 0);  // no pixel_locs: generated in the kernel
    a.q = take(q_elem_bytes(j.q_format) * np * (size_t)lay.len);
    a.qscale = take(4 * np);
    a.rot = take(8 * (size_t)lay.len);
    a.pc = take(sizeof(PulseConst) * np);
    a.hdr = take(sizeof(BpHeader));
    a.out = take(16 * npix);
    a.total = off;
    return a;
}

int check_job(const sarbp_job *j) {
    if (!j || !j->phs || !j->k_r || !j->pos || !j->out) {
        set_error("sarbp_server_submit: NULL job or buffer");
        return halide_error_code_buffer_argument_is_null;
    }
    if (j->nsamples < 2 || j->npulses < 2 || j->nu < 1 || j->nv < 1) {
        set_error("sarbp_server_submit: nsamples, npulses >= 2 and nu, nv >= 1 required");
        return halide_error_code_param_too_small;
    }
    if (j->n_fft < 2 || j->n_fft > kMaxFft || (j->n_fft & (j->n_fft - 1))) {
        set_error("sarbp_server_submit: N_fft (%d) must be a power of two in [2, %d] (longer "
                  "profiles and other lengths run through plans and backprojection())",
                  j->n_fft, kMaxFft);
        return halide_error_code_param_too_large;
    }
    if ((int64_t)j->nu * j->nv > INT32_MAX) {
        set_error("sarbp_server_submit: image too large");
        return halide_error_code_buffer_extents_too_large;
    }
    if (j->norm_r0_policy < SARBP_POLICY_DEFAULT || j->norm_r0_policy > 2 ||
        j->taylor_den_policy < SARBP_POLICY_DEFAULT || j->taylor_den_policy > 1 ||
        j->q_format < 0 || j->q_format > 1 ||
        (j->range_mode != SARBP_RANGE_F64 && j->range_mode != SARBP_RANGE_DF32)) {
        set_error("sarbp_server_submit: invalid policy or format");
        return halide_error_code_constraint_violated;
    }
    return 0;
}

bool event_done(hipEvent_t ev) { return hipEventQuery(ev) == hipSuccess; }
int advance_upload(sarbp_server *s, Job &j);
int advance_download(sarbp_server *s, Job &j);

// Enqueue the uploads of an admitted job on the stage stream: a small job's at once, then its
// ev_up; a larger job's staged arrays at once and its phase history and pixel locations by
// advance_upload.
int upload_job(sarbp_server *s, Job &j) {
    hipStream_t st = s->stage;
    const sarbp_job &d = j.d;
    const size_t np = (size_t)d.npulses, ns = (size_t)d.nsamples;
    const size_t npix = (size_t)d.nu * (size_t)d.nv;
    uint8_t *b = j.dev;
    j.stage = s->pinned.get(j.ar.host_hi - j.ar.host_lo, &j.stage_cls);
    if (!j.stage) {
        set_error("job %llu: page-locked staging allocation failed", (unsigned long long)j.id);
        return halide_error_code_out_of_memory;
    }
    uint8_t *h = (uint8_t *)j.stage - j.ar.host_lo;  // same offsets as the device block
    double *u = (double *)(h + j.ar.u), *v = (double *)(h + j.ar.v);
    const int u0 = d.nu % 2 == 0 ? -(d.nu / 2) : -(d.nu / 2) - 1;  // floor(-n / 2) (img_plane.cpp:12-14)
    const int v0 = d.nv % 2 == 0 ? -(d.nv / 2) : -(d.nv / 2) - 1;
    for (int x = 0; x < d.nu; ++x) u[x] = (double)(u0 + x) * d.d_u;
    for (int x = 0; x < d.nv; ++x) v[x] = (double)(v0 + x) * d.d_v;
    std::memcpy(h + j.ar.k_r, d.k_r, 4 * ns);
    std::memcpy(h + j.ar.pos, d.pos, 12 * np);
    std::memcpy(h + j.ar.ws, j.ws.data(), 8 * ns);
    std::memcpy(h + j.ar.wp, j.wp.data(), 8 * np);
    hipError_t e = hipMemcpyAsync(b + j.ar.host_lo, j.stage, j.ar.host_hi - j.ar.host_lo,
                                  hipMemcpyHostToDevice, st);
    j.up_total = 8 * np * ns + (d.pixel_locs ? 24 * npix : 0);
    if (e == hipSuccess && !j.small && j.up_total > kUploadChunkBytes * kUploadChunksAhead)
        return advance_upload(s, j);
    if (e == hipSuccess) e = hipMemcpyAsync(b + j.ar.phs, d.phs, 8 * np * ns, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && d.pixel_locs)
        e = hipMemcpyAsync(b + j.ar.r, d.pixel_locs, 24 * npix, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(j.ev_up, st);
    if (e != hipSuccess) {
        set_error("job %llu: upload failed: %s", (unsigned long long)j.id, hipGetErrorString(e));
        return halide_error_code_copy_to_device_failed;
    }
    j.up_sent = j.up_total;
    j.up_recorded = true;
    return 0;
}

// A larger job's phase history, then its pixel locations, in chunks of kUploadChunkBytes on the
// stage stream, at most kUploadChunksAhead outstanding; ev_up after the last one.
int advance_upload(sarbp_server *s, Job &j) {
    hipStream_t st = s->stage;
    while (j.chunks_done < j.chunks_enqueued &&
           event_done(j.ev_chunk[j.chunks_done % kUploadChunksAhead]))
        ++j.chunks_done;
    const size_t b_phs = 8 * (size_t)j.d.npulses * (size_t)j.d.nsamples;
    hipError_t e = hipSuccess;
    while (j.up_sent < j.up_total && j.chunks_enqueued - j.chunks_done < kUploadChunksAhead) {
        hipEvent_t &ev = j.ev_chunk[j.chunks_enqueued % kUploadChunksAhead];
        if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) break;
        // a chunk stays within one of the two arrays
        const bool in_phs = j.up_sent < b_phs;
        const size_t at = in_phs ? j.up_sent : j.up_sent - b_phs;
        const size_t n = std::min(kUploadChunkBytes, (in_phs ? b_phs : j.up_total - b_phs) - at);
        uint8_t *dst = j.dev + (in_phs ? j.ar.phs : j.ar.r) + at;
        const uint8_t *src = (const uint8_t *)(in_phs ? j.d.phs : (const void *)j.d.pixel_locs) + at;
        if ((e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st)) != hipSuccess ||
            (e = hipEventRecord(ev, st)) != hipSuccess)
            break;
        j.up_sent += n;
        ++j.chunks_enqueued;
    }
    if (e == hipSuccess && j.up_sent == j.up_total && !j.up_recorded) {
        if ((e = hipEventRecord(j.ev_up, st)) == hipSuccess) j.up_recorded = true;
    }
    if (e != hipSuccess) {
        set_error("job %llu: upload failed: %s", (unsigned long long)j.id, hipGetErrorString(e));
        return halide_error_code_copy_to_device_failed;
    }
    return 0;
}

// Enqueue image plane, pulse table and range profiles of an uploaded job on `st`, and fill its
// accumulation arguments.
int prepare_job(sarbp_server *s, Job &j, hipStream_t st) {
    const sarbp_job &d = j.d;
    const size_t npix = (size_t)d.nu * (size_t)d.nv;
    uint8_t *b = j.dev;
    hipError_t e = hipSuccess;
    // without pixel_locs the accumulation generates the image plane itself (grid mode)
    const PixelGrid grid = d.pixel_locs ? PixelGrid{}
                                        : make_grid(d.nu, d.nv, d.d_u, d.d_v, d.u_hat, d.v_hat);
    const int n = d.n_fft;
    const float *kr = (const float *)(b + j.ar.k_r);
    e = launch_pulse_consts((const float *)(b + j.ar.pos), kr, d.nsamples, d.npulses, j.lay,
                            d.norm_r0_policy, grid.on ? &grid : nullptr, (PulseConst *)(b + j.ar.pc),
                            (BpHeader *)(b + j.ar.hdr), st);
    if (e == hipSuccess)
        e = launch_range_profiles(b + j.ar.phs, kr, (const double *)(b + j.ar.ws),
                                  (const double *)(b + j.ar.wp), d.nsamples, d.npulses, j.lay,
                                  s->tw[ilog2(n)], (float2 *)(b + j.ar.rot), b + j.ar.q, d.q_format,
                                  d.q_format == SARBP_Q_C16 ? (float *)(b + j.ar.qscale) : nullptr,
                                  st, kSharedFftBlocks);
    if (e != hipSuccess) {
        set_error("job %llu: range profiles failed: %s", (unsigned long long)j.id,
                  hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    BpArgs &a = j.args;
    a.q = b + j.ar.q;
    a.q_format = d.q_format;
    a.qscale = d.q_format == SARBP_Q_C16 ? (const float *)(b + j.ar.qscale) : nullptr;
    a.n_fft = n;
    a.np = d.npulses;
    a.pc = (const PulseConst *)(b + j.ar.pc);
    a.hdr = (const BpHeader *)(b + j.ar.hdr);
    a.ext = j.lay.ext;
    a.len = j.lay.len;
    a.r_cover2 = j.lay.r_cover * j.lay.r_cover;
    a.r = (const double *)(b + j.ar.r);
    a.r_min = 0;
    a.r_extent = (int64_t)npix;
    a.nu = d.nu;
    a.nv = d.nv;
    a.x0 = 0;
    a.y0 = 0;
    a.width = d.nu;
    a.height = d.nv;
    a.full_w = d.nu;
    a.full_h = d.nv;
    a.out = (double2 *)(b + j.ar.out);
    a.ppt = j.ppt;
    a.g = grid;
    a.row_windows = j.row_windows;
    a.range_df = d.range_mode == SARBP_RANGE_DF32;
    j.prepared = true;
    return 0;
}

// Output download of an accumulated job and its completion event, on the job's download stream:
// one copy, or a larger job's output of more than kUploadChunksAhead chunks by advance_download
// (the copies of both
// directions share the device's copy engines, so that a small job's upload or download waits for
// at most kUploadChunksAhead chunks of a large one). On failure nothing is recorded.
int finish_job(sarbp_server *s, Job &j) {
    j.phase = Phase::Downloading;
    const size_t npix = (size_t)j.d.nu * (size_t)j.d.nv;
    if (!j.small && 16 * npix > kUploadChunkBytes * kUploadChunksAhead) {
        j.dl_sent = 0;
        j.chunks_enqueued = j.chunks_done = 0;  // the upload's chunks completed before
        return advance_download(s, j);
    }
    hipStream_t st = j.small ? s->dl_small : s->dl_big;
    hipError_t e = hipMemcpyAsync(j.d.out, j.dev + j.ar.out, 16 * npix, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && (e = hipEventRecord(j.ev_done, st)) == hipSuccess) j.done_recorded = true;
    if (e != hipSuccess) {
        set_error("job %llu: download failed: %s", (unsigned long long)j.id, hipGetErrorString(e));
        return halide_error_code_copy_to_host_failed;
    }
    return 0;
}

int advance_download(sarbp_server *s, Job &j) {
    hipStream_t st = s->dl_big;
    while (j.chunks_done < j.chunks_enqueued &&
           event_done(j.ev_chunk[j.chunks_done % kUploadChunksAhead]))
        ++j.chunks_done;
    const size_t total = 16 * (size_t)j.d.nu * (size_t)j.d.nv;
    hipError_t e = hipSuccess;
    while (j.dl_sent < total && j.chunks_enqueued - j.chunks_done < kUploadChunksAhead) {
        hipEvent_t &ev = j.ev_chunk[j.chunks_enqueued % kUploadChunksAhead];
        if (!ev && (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) break;
        const size_t n = std::min(kUploadChunkBytes, total - j.dl_sent);
        if ((e = hipMemcpyAsync((uint8_t *)j.d.out + j.dl_sent, j.dev + j.ar.out + j.dl_sent, n,
                                hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipEventRecord(ev, st)) != hipSuccess)
            break;
        j.dl_sent += n;
        ++j.chunks_enqueued;
    }
    if (e == hipSuccess && j.dl_sent == total && (e = hipEventRecord(j.ev_done, st)) == hipSuccess)
        j.done_recorded = true;
    if (e != hipSuccess) {
        set_error("job %llu: download failed: %s", (unsigned long long)j.id, hipGetErrorString(e));
        return halide_error_code_copy_to_host_failed;
    }
    return 0;
}

// A job whose enqueued work failed part way: it completes with `code` once the device is idle
// (nothing enqueued for it can still touch its arena block or staging buffer afterwards).
void fail_inflight(sarbp_server *s, Job &j, int code) {
    (void)s;
    (void)hipDeviceSynchronize();
    j.done_recorded = true;  // ev_done is not recorded: it queries as complete
    j.status = code;
}

// Row bands of a large job's accumulation. The tiles of the bands are those of one launch over
// the image, so the result is the same bit for bit. At most kBandsAhead bands are enqueued at a
// time; the scheduler adds the next band when the oldest completes, so work on any stream that
// shares a hardware queue with this job's stream waits for at most kBandsAhead bands.
int band_rows(const BpArgs &a) {
    const int64_t cols = (a.width + 15) / 16, th = tile_height(a.ppt);
    return (int)(std::max<int64_t>(1, kBandPixels / (16 * th * cols)) * th);
}

// Pixel-pulses of a banded job's rows not yet enqueued.
int64_t remaining_pp(const Job &j) {
    return (int64_t)(j.args.height - j.next_row) * j.args.width * j.d.npulses;
}

// Whether j may enqueue bands now: of the banded jobs on its stream, the one with the least work
// left (the earlier job on a tie). The others wait with at most the bands they already have in
// the queue, so jobs sharing a stream finish shortest-remaining-first rather than all late
// together; a job becomes the leader when the one before it has enqueued its last band. Jobs that
// have enqueued nothing for kBandPatienceNs come first, the longest waiting first.
bool stream_leader(const sarbp_server *s, const Job &j, int64_t now) {
    auto key = [now](const Job &x) {
        const bool waited = now - x.t_band > kBandPatienceNs;
        return std::make_tuple(waited ? 0 : 1, waited ? x.t_band : remaining_pp(x), x.id);
    };
    const auto mine = key(j);
    for (const auto &op : s->inflight) {
        const Job &o = *op;
        if (&o == &j || o.stream != j.stream || o.phase != Phase::Computing || !o.banding)
            continue;
        if (key(o) < mine) return false;
    }
    return true;
}

// Enqueue bands while fewer than kBandsAhead are outstanding (and j leads its stream); after the
// last band, ev_acc.
int advance_bands(sarbp_server *s, Job &j) {
    hipStream_t st = s->streams[j.stream];
    while (j.bands_done < j.bands_enqueued &&
           hipEventQuery(j.ev_band[j.bands_done % kBandsAhead]) == hipSuccess)
        ++j.bands_done;
    if (s->cfg.band_order == 0 && !stream_leader(s, j, now_ns(s->t0))) return 0;
    const int rows = band_rows(j.args);
    hipError_t e = hipSuccess;
    while (j.next_row < j.args.height && j.bands_enqueued - j.bands_done < kBandsAhead) {
        BpArgs a = j.args;
        a.y0 = j.args.y0 + j.next_row;
        a.height = std::min(rows, j.args.height - j.next_row);
        a.out = j.args.out + (size_t)j.next_row * (size_t)j.args.width;
        e = launch_accumulate(a, st);
        if (e == hipSuccess) e = hipEventRecord(j.ev_band[j.bands_enqueued % kBandsAhead], st);
        if (e != hipSuccess) {
            set_error("job %llu: accumulation failed: %s", (unsigned long long)j.id,
                      hipGetErrorString(e));
            return halide_error_code_device_run_failed;
        }
        j.next_row += a.height;
        ++j.bands_enqueued;
        j.t_band = now_ns(s->t0);
    }
    if (j.next_row >= j.args.height && j.banding) {
        j.banding = false;
        if ((e = hipEventRecord(j.ev_acc, st)) != hipSuccess) {
            set_error("job %llu: event record failed: %s", (unsigned long long)j.id,
                      hipGetErrorString(e));
            return halide_error_code_device_run_failed;
        }
        j.phase = Phase::Finishing;
    }
    return 0;
}

// Accumulate an uploaded job that runs alone; the first job of an undecided shape class first
// times every tile candidate over a band of its top rows, which the full pass then rewrites.
int run_alone(sarbp_server *s, Job &j) {
    hipStream_t st = s->streams[j.stream];
    Tuning &tu = s->tuning[j.cls];
    hipError_t e = hipSuccess;
    j.phase = Phase::Computing;
    if (!j.prepared) {
        int rv = prepare_job(s, j, st);
        if (rv) return rv;
    }
    if (s->cfg.autotune && tu.best == 0 && !tu.running) {
        tu.running = true;
        const int64_t row_pp = (int64_t)j.d.nu * j.d.npulses;
        const int band = (int)std::min<int64_t>(
            j.d.nv, std::max<int64_t>(192, (4000000000LL / row_pp) / 192 * 192));
        // each candidate twice, in mirrored order, so that work on other streams that overlaps one
        // probe does not decide the class; the faster of its two times counts
        constexpr int W = kTuneWindows;
        for (int cand : {4, 3, 2, 4 | W, 3 | W, 3 | W, 4 | W, 2, 3, 4}) {
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            BpArgs args = j.args;
            args.ppt = cand & (W - 1);
            args.row_windows = (cand & W) != 0;
            args.height = band;
            (void)hipEventRecord(a, st);
            e = launch_accumulate(args, st);
            (void)hipEventRecord(b, st);
            tu.probes.push_back({cand, {a, b}});
            if (e != hipSuccess) break;
        }
        j.ppt = default_tile_ppt(j.args.width, j.args.height);
        j.row_windows = 0;
    }
    if (e != hipSuccess) {
        set_error("job %llu: accumulation failed: %s", (unsigned long long)j.id,
                  hipGetErrorString(e));
        return halide_error_code_device_run_failed;
    }
    j.args.ppt = j.ppt;
    j.args.row_windows = j.row_windows;
    // bands follow the range profiles on the job's stream (a small job sent here from the fused
    // path was prepared before)
    j.t_launch = now_ns(s->t0);
    if (!j.t_ready) j.t_ready = j.t_launch;
    j.banding = true;
    j.t_band = j.t_launch;
    j.next_row = 0;
    for (int k = 0; k < kBandsAhead; ++k)
        if (!j.ev_band[k] && hipEventCreateWithFlags(&j.ev_band[k], hipEventDisableTiming) != hipSuccess) {
            set_error("job %llu: event creation failed", (unsigned long long)j.id);
            return halide_error_code_device_run_failed;
        }
    return advance_bands(s, j);
}

size_t tiles_of(const Job &j, int th) {
    return (size_t)((j.d.nv + th - 1) / th) * (size_t)((j.d.nu + 15) / 16);
}

// One accumulation launch over the tiles of `group` (prepared jobs of one format and tile shape,
// table within the slot), using the free launch-table slot `slot`; the jobs are downloaded once
// the scheduler sees the slot's event complete. On failure every job of the group is failed.
void run_fused(sarbp_server *s, FuseSlot &slot, std::vector<Job *> &group) {
    const int ppt = group[0]->ppt;
    const int th = tile_height(ppt);
    size_t ntiles = 0;
    for (Job *j : group) ntiles += tiles_of(*j, th);
    const size_t ba = align256(sizeof(BpArgs) * group.size());
    BpArgs *args = (BpArgs *)slot.host;
    FusedTile *tiles = (FusedTile *)((uint8_t *)slot.host + ba);
    size_t t = 0;
    for (size_t k = 0; k < group.size(); ++k) {
        const Job *j = group[k];
        args[k] = j->args;
        args[k].cubic = grid_cubic(args[k]);
        for (int by = 0; by < j->d.nv; by += th)
            for (int bx = 0; bx < j->d.nu; bx += 16) tiles[t++] = {(int)k, bx, by, 0};
    }
    hipStream_t fs = s->fuse_stream;
    hipError_t e = hipMemcpyAsync(slot.dev, slot.host, ba + sizeof(FusedTile) * ntiles,
                                  hipMemcpyHostToDevice, fs);
    if (e == hipSuccess)
        e = launch_accumulate_fused((const BpArgs *)slot.dev,
                                    (const FusedTile *)((uint8_t *)slot.dev + ba), (int)ntiles,
                                    group[0]->d.q_format, ppt, group[0]->args.g.on != 0, fs);
    if (e == hipSuccess) e = hipEventRecord(slot.ev, fs);
    if (e != hipSuccess) {
        set_error("fused accumulation failed: %s", hipGetErrorString(e));
        for (Job *j : group) fail_inflight(s, *j, halide_error_code_device_run_failed);
        return;
    }
    slot.busy = true;
    slot.group = group;
    const int64_t t_launch = now_ns(s->t0);
    for (Job *j : group) {
        j->fused = true;
        j->phase = Phase::Computing;
        j->t_launch = t_launch;
    }
}

void retire(sarbp_server *s, Job &j) {
    s->arena.release(j.off, j.ar.total);
    s->used_bytes -= (int64_t)j.ar.total;
    if (!j.small) s->stream_load[j.stream] -= j.pp;
    s->pinned.put(j.stage, j.stage_cls);
    j.stage = nullptr;
    (void)hipEventDestroy(j.ev_done);
    (void)hipEventDestroy(j.ev_up);
    (void)hipEventDestroy(j.ev_ready);
    (void)hipEventDestroy(j.ev_acc);
    (void)hipEventDestroy(j.ev_begin);
    for (hipEvent_t &ev : j.ev_band)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t &ev : j.ev_chunk)
        if (ev) (void)hipEventDestroy(ev);
}

// Pulse table and range profiles of an uploaded small job on the prep stream, then ev_ready.
int prep_small(sarbp_server *s, Job &j) {
    int rv = prepare_job(s, j, s->prep);
    if (rv) return rv;
    if (hipEventRecord(j.ev_ready, s->prep) != hipSuccess) {
        set_error("job %llu: event record failed", (unsigned long long)j.id);
        return halide_error_code_device_run_failed;
    }
    j.phase = Phase::Prepping;
    return 0;
}

void scheduler(sarbp_server *s) {
    (void)hipSetDevice(s->cfg.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        // 1. completions, and the step after each completed step
        bool progressed = false;
        for (size_t i = 0; i < s->inflight.size();) {
            Job &j = *s->inflight[i];
            if (j.done_recorded && event_done(j.ev_done)) {
                j.t_end = now_ns(s->t0);
                if (j.status > 0) j.status = 0;
                (void)hipEventElapsedTime(&j.device_ms, j.ev_begin, j.ev_done);
                retire(s, j);
                s->inflight.erase(s->inflight.begin() + (long)i);
                progressed = true;
            } else {
                ++i;
            }
        }
        for (FuseSlot &sl : s->slots) {
            if (!sl.busy || !event_done(sl.ev)) continue;
            sl.busy = false;
            const int64_t t_acc = now_ns(s->t0);
            for (Job *j : sl.group) {
                j->t_acc = t_acc;
                int rv = finish_job(s, *j);
                if (rv) fail_inflight(s, *j, rv);
            }
            sl.group.clear();
        }
        for (auto &jp : s->inflight) {
            Job &j = *jp;
            if (j.done_recorded) continue;
            int rv = 0;
            if (j.phase == Phase::Downloading) {
                rv = advance_download(s, j);  // a larger job's output, chunk by chunk
            } else if (j.phase == Phase::Uploading && !j.up_recorded) {
                rv = advance_upload(s, j);
            } else if (j.phase == Phase::Uploading && event_done(j.ev_up)) {
                j.t_up = now_ns(s->t0);
                rv = j.small ? prep_small(s, j) : run_alone(s, j);
            } else if (j.phase == Phase::Prepping && event_done(j.ev_ready)) {
                j.t_ready = now_ns(s->t0);
                j.phase = Phase::Ready;
                s->fuse_pending.push_back(&j);
            } else if (j.phase == Phase::Computing && j.banding) {
                rv = advance_bands(s, j);
                if (rv) j.banding = false;
            }
            if (rv == 0 && j.phase == Phase::Finishing && event_done(j.ev_acc)) {
                j.t_acc = now_ns(s->t0);
                rv = finish_job(s, j);
            }
            if (rv) fail_inflight(s, j, rv);
        }
        for (auto &kv : s->tuning) {
            Tuning &tu = kv.second;
            if (!tu.running) continue;
            bool done = true;
            for (auto &pr : tu.probes) done = done && event_done(pr.second.second);
            if (!done) continue;
            std::map<int, float> fastest;  // candidate -> its faster probe
            for (auto &pr : tu.probes) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, pr.second.first, pr.second.second);
                auto it = fastest.find(pr.first);
                if (it == fastest.end() || ms < it->second) fastest[pr.first] = ms;
                (void)hipEventDestroy(pr.second.first);
                (void)hipEventDestroy(pr.second.second);
            }
            float best_ms = 1e30f;
            for (auto &kv2 : fastest)
                if (kv2.second < best_ms) {
                    best_ms = kv2.second;
                    tu.best = kv2.first;
                }
            tu.probes.clear();
            tu.running = false;
        }
        if (progressed) s->cv_done.notify_all();

        // 2. admission in arrival order; jobs that fit may pass a head that does not, until the
        // head has waited kHeadPatienceNs
        const int64_t now = now_ns(s->t0);
        for (auto it = s->queue.begin(); it != s->queue.end();) {
            if ((int)s->inflight.size() >= s->cfg.max_inflight) break;
            Job &j = **it;
            const size_t off = s->arena.alloc(j.ar.total);
            if (off == SIZE_MAX) {
                if (it == s->queue.begin() && now - j.t_submit > kHeadPatienceNs) break;
                ++it;
                continue;
            }
            auto sp = *it;
            it = s->queue.erase(it);
            const int64_t h0 = now_ns(s->t0);
            j.t_admit = h0;
            j.status = 2;
            j.off = off;
            j.dev = s->arena_base + off;
            s->used_bytes += (int64_t)j.ar.total;
            Tuning &tu = s->tuning[j.cls];
            j.ppt = tu.best ? tu.best & (kTuneWindows - 1) : default_tile_ppt(j.d.nu, j.d.nv);
            j.row_windows = (tu.best & kTuneWindows) != 0;
            // a small job of an untuned class runs alone once, on a big stream, to be timed
            const bool fuse = j.small && (tu.best != 0 || !s->cfg.autotune);
            j.small = fuse;
            j.stream = (int)(std::min_element(s->stream_load.begin(), s->stream_load.end()) -
                             s->stream_load.begin());
            if (!j.small) s->stream_load[j.stream] += j.pp;
            (void)hipEventCreateWithFlags(&j.ev_up, hipEventDisableTiming);
            (void)hipEventCreateWithFlags(&j.ev_ready, hipEventDisableTiming);
            (void)hipEventCreateWithFlags(&j.ev_acc, hipEventDisableTiming);
            (void)hipEventCreate(&j.ev_done);
            (void)hipEventCreate(&j.ev_begin);
            (void)hipEventRecord(j.ev_begin, s->stage);
            s->inflight.push_back(sp);
            j.phase = Phase::Uploading;
            int rv = upload_job(s, j);
            if (rv != 0) fail_inflight(s, j, rv);
            j.host_ns = now_ns(s->t0) - h0;
        }

        // 3. small jobs: while every launch-table slot is in flight, prepared jobs wait and gather;
        // when a slot is free, the waiting jobs go into one launch per (format, tile shape, grid
        // mode), as many as the slot's table and fuse_launch_pixel_pulses allow. An idle GPU thus starts a
        // small job at once and a busy one receives them in large batches.
        for (FuseSlot &sl : s->slots) {
            if (s->fuse_pending.empty()) break;
            if (sl.busy) continue;
            const int fmt = s->fuse_pending[0]->d.q_format, ppt = s->fuse_pending[0]->ppt;
            const int grid = s->fuse_pending[0]->args.g.on;
            const int th = tile_height(ppt);
            std::vector<Job *> group, rest;
            size_t bytes = 0;
            int64_t pp = 0;
            for (Job *j : s->fuse_pending) {
                const size_t need = sizeof(BpArgs) + 256 + sizeof(FusedTile) * tiles_of(*j, th);
                if (j->d.q_format == fmt && j->ppt == ppt && j->args.g.on == grid &&
                    (group.empty() ||
                     (bytes + need <= kFuseSlotBytes && pp + j->pp <= s->cfg.fuse_launch_pixel_pulses))) {
                    group.push_back(j);
                    bytes += need;
                    pp += j->pp;
                } else {
                    rest.push_back(j);
                }
            }
            s->fuse_pending.swap(rest);
            if (bytes > kFuseSlotBytes) {  // one job larger than a table: accumulate it alone
                Job &j = *group[0];
                j.small = false;  // downloads on dl_big; its stream load is not counted
                s->stream_load[j.stream] += j.pp;
                int rv = run_alone(s, j);
                if (rv) fail_inflight(s, j, rv);
                continue;
            }
            run_fused(s, sl, group);
        }
        if (s->stop && s->queue.empty() && s->inflight.empty()) break;
        if (s->inflight.empty() && s->queue.empty()) {
            s->cv_work.wait(lk, [s] { return s->stop || !s->queue.empty(); });
        } else {
            // poll device progress at 20 microsecond granularity, releasing the lock meanwhile
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

void destroy_resources(sarbp_server *s) {
    for (auto st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    s->streams.clear();
    for (hipStream_t *st : {&s->stage, &s->prep, &s->fuse_stream, &s->dl_small, &s->dl_big}) {
        if (!*st) continue;
        (void)hipStreamSynchronize(*st);
        (void)hipStreamDestroy(*st);
        *st = nullptr;
    }
    for (FuseSlot &sl : s->slots) {
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.dev) (void)hipFree(sl.dev);
        if (sl.host) (void)hipHostFree(sl.host);
        sl = FuseSlot{};
    }
    if (s->arena_base) (void)hipFree(s->arena_base);
    s->arena_base = nullptr;
}

}  // namespace

extern "C" void sarbp_server_config_init(sarbp_server_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->device = 0;
    c->max_inflight = 64;
    c->hbm_budget_bytes = (int64_t)64 << 30;
    c->fuse_max_pixel_pulses = (int64_t)1 << 31;
    c->num_streams = 2;  // streams for large jobs, besides the stage, prep, fuse and download ones
    c->autotune = 1;
    c->fuse_launch_pixel_pulses = kFusePixelPulses;
}

extern "C" int sarbp_server_create(const sarbp_server_config *cfg, sarbp_server **out) {
    if (!cfg || !out) {
        set_error("sarbp_server_create: NULL argument");
        return halide_error_code_buffer_argument_is_null;
    }
    *out = nullptr;
    if (cfg->max_inflight < 1 || cfg->num_streams < 1 || cfg->num_streams > 16 ||
        cfg->hbm_budget_bytes <= 0 || cfg->band_order < 0 || cfg->band_order > 1) {
        set_error("sarbp_server_create: max_inflight >= 1, 1 <= num_streams <= 16, a positive "
                  "budget and band_order 0 or 1 are required");
        return halide_error_code_constraint_violated;
    }
    if (sarbp_device_count() <= cfg->device || cfg->device < 0) {
        set_error("sarbp_server_create: device %d is not visible", cfg->device);
        return halide_error_code_no_device_interface;
    }
    auto s = std::make_unique<sarbp_server>();
    s->cfg = *cfg;
    if (s->cfg.fuse_launch_pixel_pulses <= 0) s->cfg.fuse_launch_pixel_pulses = kFusePixelPulses;
    s->stream_load.assign((size_t)cfg->num_streams, 0);
    s->t0 = std::chrono::steady_clock::now();
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(cfg->device);
    int rv = 0;
    if (hipMalloc((void **)&s->arena_base, (size_t)cfg->hbm_budget_bytes) != hipSuccess) {
        set_error("sarbp_server_create: the HBM budget (%lld bytes) cannot be allocated",
                  (long long)cfg->hbm_budget_bytes);
        s->arena_base = nullptr;
        (void)hipGetLastError();
        rv = halide_error_code_device_malloc_failed;
    }
    s->arena.reset((size_t)cfg->hbm_budget_bytes);
    for (int i = 0; rv == 0 && i < cfg->num_streams + 5; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            set_error("sarbp_server_create: stream creation failed");
            rv = halide_error_code_device_run_failed;
            break;
        }
        if (i < cfg->num_streams) s->streams.push_back(st);
        else if (i == cfg->num_streams) s->stage = st;
        else if (i == cfg->num_streams + 1) s->prep = st;
        else if (i == cfg->num_streams + 2) s->fuse_stream = st;
        else if (i == cfg->num_streams + 3) s->dl_small = st;
        else s->dl_big = st;
    }
    for (FuseSlot &sl : s->slots) {
        if (rv) break;
        if (hipHostMalloc(&sl.host, kFuseSlotBytes, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&sl.dev, kFuseSlotBytes) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            set_error("sarbp_server_create: launch table allocation failed");
            rv = halide_error_code_device_malloc_failed;
        }
    }
    for (int lg = 1; rv == 0 && (1 << lg) <= kMaxFft; ++lg) {
        s->tw[lg] = twiddles(cfg->device, 1 << lg, nullptr);
        if (!s->tw[lg]) {
            set_error("sarbp_server_create: twiddle table allocation failed");
            rv = halide_error_code_device_malloc_failed;
        }
    }
    if (rv) {
        destroy_resources(s.get());
        (void)hipSetDevice(prev);
        return rv;
    }
    (void)hipSetDevice(prev);
    sarbp_server *raw = s.release();
    raw->worker = std::thread(scheduler, raw);
    *out = raw;
    return 0;
}

extern "C" int sarbp_server_submit(sarbp_server *s, const sarbp_job *job, uint64_t *job_id) {
    if (!s) {
        set_error("sarbp_server_submit: NULL server");
        return halide_error_code_buffer_argument_is_null;
    }
    int rv = check_job(job);
    if (rv) return rv;
    auto j = std::make_shared<Job>();
    j->d = *job;
    // SARBP_POLICY_DEFAULT: the process-wide defaults at submission, as a plan would take them
    int def_norm, def_den;
    sarbp_get_default_policies(&def_norm, &def_den);
    if (j->d.norm_r0_policy == SARBP_POLICY_DEFAULT) j->d.norm_r0_policy = def_norm;
    if (j->d.taylor_den_policy == SARBP_POLICY_DEFAULT) j->d.taylor_den_policy = def_den;
    job = &j->d;
    j->lay = row_layout(job->nsamples, job->n_fft, job->delta_r);
    j->ar = layout_of(*job, j->lay);
    if ((int64_t)j->ar.total > s->cfg.hbm_budget_bytes) {
        set_error("sarbp_server_submit: job needs %zu bytes, more than the HBM budget (%lld)",
                  j->ar.total, (long long)s->cfg.hbm_budget_bytes);
        return halide_error_code_buffer_allocation_too_large;
    }
    j->ws.resize((size_t)job->nsamples);
    j->wp.resize((size_t)job->npulses);
    taylor_window(job->nsamples, job->taylor_s_l, job->taylor_den_policy, j->ws.data());
    taylor_window(job->npulses, job->taylor_s_l, job->taylor_den_policy, j->wp.data());
    const int64_t pp = (int64_t)job->nu * job->nv * job->npulses;
    j->pp = pp;
    j->small = s->cfg.fuse_max_pixel_pulses > 0 && pp <= s->cfg.fuse_max_pixel_pulses;
    j->cls = std::make_tuple(job->n_fft, job->q_format, ilog2(job->npulses),
                             ilog2((int64_t)job->nu * job->nv));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->id = s->next_id++;
        j->t_submit = now_ns(s->t0);
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        if (job_id) *job_id = j->id;
    }
    s->cv_work.notify_one();
    return 0;
}

extern "C" int sarbp_server_wait(sarbp_server *s, uint64_t job_id, int32_t timeout_ms,
                                 sarbp_job_stats *stats) {
    if (!s) {
        set_error("sarbp_server_wait: NULL server");
        return halide_error_code_buffer_argument_is_null;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) {
        set_error("sarbp_server_wait: unknown job %llu", (unsigned long long)job_id);
        return halide_error_code_generic_error;
    }
    std::shared_ptr<Job> j = it->second;
    auto done = [&] { return j->status <= 0; };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return 1;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->t_submit_ns = j->t_submit;
        stats->t_admit_ns = j->t_admit;
        stats->t_end_ns = j->t_end;
        stats->device_bytes = (int64_t)j->ar.total;
        stats->fused = j->fused ? 1 : 0;
        stats->tile_ppt = j->ppt;
        stats->row_windows = j->row_windows;
        stats->status = j->status;
        stats->host_enqueue_us = (int32_t)std::min<int64_t>(j->host_ns / 1000, INT32_MAX);
        stats->device_us = (int32_t)std::min<double>(1000.0 * j->device_ms, (double)INT32_MAX);
        auto since_admit = [&](int64_t t) {
            return t ? (int32_t)std::min<int64_t>((t - j->t_admit) / 1000, INT32_MAX) : -1;
        };
        stats->up_us = since_admit(j->t_up);
        stats->ready_us = since_admit(j->t_ready);
        stats->launch_us = since_admit(j->t_launch);
        stats->acc_us = since_admit(j->t_acc);
    }
    s->jobs.erase(job_id);
    if (j->status < 0) set_error("job %llu failed (%d)", (unsigned long long)job_id, j->status);
    return j->status < 0 ? j->status : 0;
}

extern "C" int sarbp_server_destroy(sarbp_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(s->cfg.device);
    destroy_resources(s);
    (void)hipSetDevice(prev);
    delete s;
    return 0;
}
