// Kernel timing: event pairs around launches, summed per kernel kind.
#include <cstring>

#include "mqr_internal.h"

namespace mqr {

thread_local Profiler *t_prof = nullptr;

const char *kernel_kind_name(int k) {
    static const char *names[K_COUNT] = {"panel_gemm", "vta_gemm", "panel_small", "w2",
                                         "update_gemm", "trsm", "gemm", "resid",
                                         "cast", "other", "lookahead", "apply_q",
                                         "near_update"};
    return (k >= 0 && k < K_COUNT) ? names[k] : "?";
}

ProfScope::ProfScope(int kind, double flops, double bytes, hipStream_t stream)
    : p(t_prof), idx(0), s(stream) {
    if (!p || !p->on || !((p->mask >> kind) & 1u)) {
        p = nullptr;
        return;
    }
    if (p->used + 2 > p->pool.size()) {
        for (int i = 0; i < 256; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) {
                p = nullptr;
                return;
            }
            p->pool.push_back(e);
        }
    }
    Profiler::Rec r;
    r.kind = kind;
    r.flops = flops;
    r.bytes = bytes;
    r.a = p->pool[p->used++];
    r.b = p->pool[p->used++];
    (void)hipEventRecord(r.a, s);
    idx = p->recs.size();
    p->recs.push_back(r);
}

ProfScope::~ProfScope() {
    if (p) (void)hipEventRecord(p->recs[idx].b, s);
}

}  // namespace mqr

using namespace mqr;

extern "C" {

int mqr_profile_enable(mqr_handle_t h, int on) {
    if (!h) return -1;
    h->prof.on = on != 0;
    h->prof.recs.clear();
    h->prof.used = 0;
    return MQR_SUCCESS;
}

int mqr_profile_set_mask(mqr_handle_t h, unsigned mask) {
    if (!h) return -1;
    h->prof.mask = mask;
    return MQR_SUCCESS;
}

int mqr_profile_read(mqr_handle_t h, mqr_kernel_stat_t *stats, int max_stats, int *count) {
    if (!h) return -1;
    if (max_stats > 0 && !stats) return -2;
    if (!count) return -4;
    MQR_HIP(hipStreamSynchronize(h->stream));
    MQR_HIP(hipStreamSynchronize(h->side_stream));
    MQR_HIP(hipStreamSynchronize(h->far_stream));
    mqr_kernel_stat_t acc[K_COUNT];
    std::memset(acc, 0, sizeof(acc));
    for (const auto &r : h->prof.recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        mqr_kernel_stat_t &s = acc[r.kind];
        s.launches += 1;
        s.total_ms += ms;
        s.flops += r.flops;
        s.bytes += r.bytes;
    }
    int c = 0;
    for (int k = 0; k < K_COUNT; ++k) {
        if (acc[k].launches == 0) continue;
        std::strncpy(acc[k].name, kernel_kind_name(k), sizeof(acc[k].name) - 1);
        if (c < max_stats) stats[c] = acc[k];
        ++c;
    }
    *count = c;
    h->prof.recs.clear();
    h->prof.used = 0;
    return MQR_SUCCESS;
}

}  // extern "C"
