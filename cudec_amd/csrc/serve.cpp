/* Serving loop: continuous batching of LZ4 decode requests over the tuned
 * batch decoder, with first-come-first-served admission against an HBM
 * output budget (the contract is documented in cudec_mi355x.h).
 *
 * One host thread runs the loop. Each pass it (1) moves requests whose
 * scheduled arrival has passed into the waiting queue, (2) polls the
 * in-flight launches and retires every request of a finished one (results
 * and digests are already in pinned host memory by then; its output region
 * goes back to the pool), (3) admits waiting requests in arrival order while
 * their output fits in the pool, and (4) launches every admitted request as
 * one batch on an idle stream. Up to max_inflight launches run at once on
 * separate streams, so requests that arrive while others are on the device
 * start without waiting for them. */
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "cudec_mi355x.h"
#include "hip_owner.h"

namespace cudec_amd {
namespace {

constexpr size_t kOutAlign = 256;
constexpr int kMaxInflightLimit = 16;

/* First-fit allocator over [0, size) that merges neighbouring free ranges. */
class OutputPool {
public:
    explicit OutputPool(size_t size) {
        if (size != 0) {
            free_[0] = size;
        }
    }
    bool take(size_t n, size_t* off) {
        if (n == 0) {
            /* a request of zero-capacity chunks holds no pool bytes: it is
             * always admissible, even with a zero budget or a full pool */
            *off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                const size_t o = it->first, rest = it->second - n;
                free_.erase(it);
                if (rest != 0) {
                    free_[o + n] = rest;
                }
                *off = o;
                return true;
            }
        }
        return false;
    }
    void give(size_t off, size_t n) {
        if (n == 0) {
            return;
        }
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<size_t, size_t> free_;
};

/* Tables of one launch: source pointers, source sizes, destination
 * pointers, capacities (8 bytes each per chunk), in that order. */
constexpr size_t kTableBytes = 4 * sizeof(uint64_t);

struct Slot {
    Stream stream;
    Event done;
    PinnedRegion h_tables;
    DeviceRegion d_tables;
    PinnedRegion h_results;
    DeviceRegion d_results;
    PinnedRegion h_digests;
    DeviceRegion d_digests;
    std::vector<size_t> requests;
    bool busy = false;
};

struct Request {
    size_t bytes = 0;   /* output bytes held in the pool */
    size_t offset = 0;  /* pool offset of its output */
};

class Server {
public:
    Server(const void* const* src, const size_t* sizes, const size_t* caps,
           const cudec_mi355x_request* reqs, size_t nreq, size_t budget, int inflight,
           cudec_chunk_result* results, uint64_t* digests, double* latency)
        : src_(src), sizes_(sizes), caps_(caps), reqs_(reqs), nreq_(nreq), budget_(budget),
          inflight_(inflight), results_(results), digests_(digests), latency_(latency),
          pool_(budget), state_(nreq) {}

    cudec_status run(cudec_mi355x_serve_stats* stats) {
        if (pool_mem_.reserve(budget_) != hipSuccess) {
            return CUDEC_ERR_CUDA;
        }
        slots_.reserve((size_t)inflight_);
        for (int i = 0; i < inflight_; i++) {
            slots_.push_back(std::make_unique<Slot>());
            if (make_stream(slots_.back()->stream) != hipSuccess ||
                make_sync_event(slots_.back()->done) != hipSuccess) {
                return CUDEC_ERR_CUDA;
            }
        }
        for (size_t r = 0; r < nreq_; r++) {
            size_t b = 0;
            for (size_t k = 0; k < reqs_[r].chunk_count; k++) {
                b += round_up(caps_[reqs_[r].first_chunk + k]);
            }
            state_[r].bytes = b;
        }
        const auto t0 = std::chrono::steady_clock::now();
        auto now_s = [&] {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        };
        size_t next = 0, retired = 0, held = 0;
        std::deque<size_t> waiting;
        std::vector<size_t> admitted;
        double last_done = 0;
        while (retired < nreq_) {
            bool progress = false;
            const double now = now_s();
            while (next < nreq_ && reqs_[next].arrival_s <= now) {
                waiting.push_back(next++);
                progress = true;
            }
            for (auto& s : slots_) {
                if (!s->busy) {
                    continue;
                }
                const hipError_t q = hipEventQuery(s->done.get());
                if (q == hipErrorNotReady) {
                    continue;
                }
                if (q != hipSuccess) {
                    return fail();
                }
                const double t = now_s();
                retire(*s, t, &held);
                retired += s->requests.size();
                s->requests.clear();
                s->busy = false;
                last_done = t;
                progress = true;
            }
            while (!waiting.empty()) {
                const size_t r = waiting.front();
                if (!pool_.take(state_[r].bytes, &state_[r].offset)) {
                    break;  /* first come, first served: later ones wait too */
                }
                held += state_[r].bytes;
                peak_pool_ = std::max<uint64_t>(peak_pool_, held);
                admitted.push_back(r);
                waiting.pop_front();
                progress = true;
            }
            if (!admitted.empty()) {
                for (auto& s : slots_) {
                    if (s->busy) {
                        continue;
                    }
                    if (!launch(*s, admitted)) {
                        return fail();
                    }
                    admitted.clear();
                    progress = true;
                    break;
                }
            }
            if (!progress) {
                std::this_thread::yield();
            }
        }
        if (stats != nullptr) {
            stats->elapsed_s = last_done - (nreq_ != 0 ? reqs_[0].arrival_s : 0.0);
            stats->launches = launches_;
            stats->max_inflight = max_inflight_;
            stats->max_batch_requests = max_batch_requests_;
            stats->peak_pool_bytes = peak_pool_;
            stats->output_bytes = output_bytes_;
            stats->compressed_bytes = compressed_bytes_;
        }
        return CUDEC_OK;
    }

private:
    static size_t round_up(size_t v) { return (v + kOutAlign - 1) & ~(kOutAlign - 1); }

    cudec_status fail() {
        for (auto& s : slots_) {
            if (s->busy) {
                (void)hipStreamSynchronize(s->stream.get());
            }
        }
        return CUDEC_ERR_CUDA;
    }

    bool launch(Slot& s, const std::vector<size_t>& batch) {
        size_t n = 0;
        for (size_t r : batch) {
            n += reqs_[r].chunk_count;
        }
        const size_t want = n * kTableBytes;
        if (s.h_tables.capacity() < want) {
            /* grow by doubling so a growing load settles after a few launches */
            const size_t grow = std::max(want, 2 * s.h_tables.capacity());
            if (s.h_tables.reserve(grow) != hipSuccess || s.d_tables.reserve(grow) != hipSuccess) {
                return false;
            }
        }
        const size_t rb = n * sizeof(cudec_chunk_result);
        if (s.h_results.capacity() < rb) {
            const size_t grow = std::max(rb, 2 * s.h_results.capacity());
            if (s.h_results.reserve(grow) != hipSuccess || s.d_results.reserve(grow) != hipSuccess) {
                return false;
            }
        }
        if (digests_ != nullptr && s.h_digests.capacity() < n * sizeof(uint64_t)) {
            const size_t grow = std::max(n * sizeof(uint64_t), 2 * s.h_digests.capacity());
            if (s.h_digests.reserve(grow) != hipSuccess || s.d_digests.reserve(grow) != hipSuccess) {
                return false;
            }
        }
        uint64_t* t = static_cast<uint64_t*>(s.h_tables.data());
        uint64_t* t_src = t;
        uint64_t* t_size = t + n;
        uint64_t* t_dst = t + 2 * n;
        uint64_t* t_cap = t + 3 * n;
        const uintptr_t base = reinterpret_cast<uintptr_t>(pool_mem_.data());
        size_t i = 0;
        for (size_t r : batch) {
            size_t off = state_[r].offset;
            for (size_t k = 0; k < reqs_[r].chunk_count; k++, i++) {
                const size_t c = reqs_[r].first_chunk + k;
                t_src[i] = reinterpret_cast<uintptr_t>(src_[c]);
                t_size[i] = sizes_[c];
                t_dst[i] = base + off;
                t_cap[i] = caps_[c];
                off += round_up(caps_[c]);
                compressed_bytes_ += sizes_[c];
            }
        }
        const hipStream_t st = s.stream.get();
        uint64_t* d = static_cast<uint64_t*>(s.d_tables.data());
        if (hipMemcpyAsync(d, t, want, hipMemcpyHostToDevice, st) != hipSuccess) {
            return false;
        }
        auto* dres = static_cast<cudec_chunk_result*>(s.d_results.data());
        const auto cs = reinterpret_cast<cudec_stream_t>(st);
        if (cudec_lz4_decompress_batch(reinterpret_cast<const void* const*>(d), d + n,
                                       reinterpret_cast<void* const*>(d + 2 * n), d + 3 * n, n,
                                       dres, cs) != CUDEC_OK) {
            return false;
        }
        if (digests_ != nullptr) {
            if (cudec_mi355x_digest_batch(reinterpret_cast<const void* const*>(d + 2 * n), nullptr,
                                          dres, n, static_cast<uint64_t*>(s.d_digests.data()),
                                          cs) != CUDEC_OK ||
                hipMemcpyAsync(s.h_digests.data(), s.d_digests.data(), n * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, st) != hipSuccess) {
                return false;
            }
        }
        if (hipMemcpyAsync(s.h_results.data(), dres, rb, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipEventRecord(s.done.get(), st) != hipSuccess) {
            return false;
        }
        s.requests = batch;
        s.busy = true;
        launches_++;
        max_batch_requests_ = std::max<uint64_t>(max_batch_requests_, batch.size());
        uint64_t busy = 0;
        for (auto& o : slots_) {
            busy += o->busy ? 1 : 0;
        }
        max_inflight_ = std::max(max_inflight_, busy);
        return true;
    }

    void retire(Slot& s, double t, size_t* held) {
        const auto* res = static_cast<const cudec_chunk_result*>(s.h_results.data());
        const auto* dig = static_cast<const uint64_t*>(s.h_digests.data());
        size_t i = 0;
        for (size_t r : s.requests) {
            latency_[r] = t - reqs_[r].arrival_s;
            for (size_t k = 0; k < reqs_[r].chunk_count; k++, i++) {
                const size_t c = reqs_[r].first_chunk + k;
                results_[c] = res[i];
                output_bytes_ += res[i].bytes_written;
                if (digests_ != nullptr) {
                    digests_[c] = dig[i];
                }
            }
            pool_.give(state_[r].offset, state_[r].bytes);
            *held -= state_[r].bytes;
        }
    }

    const void* const* src_;
    const size_t* sizes_;
    const size_t* caps_;
    const cudec_mi355x_request* reqs_;
    size_t nreq_;
    size_t budget_;
    int inflight_;
    cudec_chunk_result* results_;
    uint64_t* digests_;
    double* latency_;
    OutputPool pool_;
    DeviceRegion pool_mem_;
    std::vector<Request> state_;
    std::vector<std::unique_ptr<Slot>> slots_;
    uint64_t launches_ = 0, max_inflight_ = 0, max_batch_requests_ = 0, peak_pool_ = 0;
    uint64_t output_bytes_ = 0, compressed_bytes_ = 0;
};

cudec_status validate(const void* const* src, const size_t* sizes, const size_t* caps,
                      size_t nchunks, const cudec_mi355x_request* reqs, size_t nreq,
                      size_t budget, int inflight, const cudec_chunk_result* results,
                      const double* latency) {
    if (src == nullptr || sizes == nullptr || caps == nullptr || reqs == nullptr ||
        results == nullptr || latency == nullptr || nchunks == 0 || nreq == 0 ||
        inflight < 1 || inflight > kMaxInflightLimit) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    double prev = -std::numeric_limits<double>::infinity();
    for (size_t r = 0; r < nreq; r++) {
        const cudec_mi355x_request& q = reqs[r];
        if (!(q.arrival_s >= prev) || q.chunk_count == 0 || q.first_chunk > nchunks ||
            q.chunk_count > nchunks - q.first_chunk) {
            return CUDEC_ERR_INVALID_ARGUMENT;
        }
        prev = q.arrival_s;
        size_t bytes = 0;
        for (size_t k = 0; k < q.chunk_count; k++) {
            const size_t cap = caps[q.first_chunk + k];
            if (cap > budget || cap > SIZE_MAX - kOutAlign) {
                return CUDEC_ERR_INVALID_ARGUMENT;
            }
            const size_t b = (cap + kOutAlign - 1) & ~(kOutAlign - 1);
            if (b > budget - bytes) {
                return CUDEC_ERR_INVALID_ARGUMENT;
            }
            bytes += b;
        }
    }
    return CUDEC_OK;
}

}  // namespace
}  // namespace cudec_amd

extern "C" cudec_status cudec_mi355x_serve_lz4(
    const void* const* h_src_ptrs, const size_t* h_src_sizes, const size_t* h_dst_caps,
    size_t chunk_count, const cudec_mi355x_request* h_requests, size_t request_count,
    size_t hbm_budget, int max_inflight, cudec_chunk_result* h_results, uint64_t* h_digests,
    double* h_latency_s, cudec_mi355x_serve_stats* stats) {
    using namespace cudec_amd;
    const cudec_status v = validate(h_src_ptrs, h_src_sizes, h_dst_caps, chunk_count, h_requests,
                                    request_count, hbm_budget, max_inflight, h_results,
                                    h_latency_s);
    if (v != CUDEC_OK) {
        return v;
    }
    if (stats != nullptr) {
        std::memset(stats, 0, sizeof(*stats));
    }
    try {
        Server s(h_src_ptrs, h_src_sizes, h_dst_caps, h_requests, request_count, hbm_budget,
                 max_inflight, h_results, h_digests, h_latency_s);
        return s.run(stats);
    } catch (...) {
        return CUDEC_ERR_CUDA;
    }
}
