// nbx_lists.cpp -- host-only parts of the C ABI (no device code).
//
// FOF's linked-list outputs (KDTree::FOF's pHead, pNext, pTail, pLen,
// KDFOF.cxx:29-31, :52-55) built from canonical labels.  Lists run in
// tree-index (slot) order; the reference's run in its search order.
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/nbx.h"

extern "C" int nbx_host_fail(int code, const char *msg);  // sets nbx_last_error()

extern "C" int nbx_fof_lists(const int32_t *labels_by_id, const int32_t *slot_to_id, int64_t n, int64_t ngroups,
                             int32_t *head, int32_t *next, int32_t *tail, int32_t *len) {
    if (!labels_by_id || !slot_to_id || n < 0 || ngroups < 0) return nbx_host_fail(NBX_ERR_ARG, "bad argument");
    std::vector<int32_t> first((size_t)ngroups + 1, -1), last((size_t)ngroups + 1, -1), cnt((size_t)ngroups + 1, 0);
    for (int64_t s = 0; s < n; ++s) {
        const int32_t id = slot_to_id[s];
        if (id < 0 || id >= n) return nbx_host_fail(NBX_ERR_ARG, "slot_to_id out of range");
        const int32_t g = labels_by_id[id];
        if (g < 0 || g > ngroups) return nbx_host_fail(NBX_ERR_ARG, "label out of range");
        if (next) next[s] = -1;
        if (g == 0) {  // in no group: a list of its own
            if (head) head[s] = (int32_t)s;
            if (tail) tail[s] = (int32_t)s;
            continue;
        }
        if (first[g] < 0) first[g] = (int32_t)s;
        else if (next) next[last[g]] = (int32_t)s;
        last[g] = (int32_t)s;
        cnt[g]++;
    }
    for (int64_t s = 0; s < n; ++s) {
        const int32_t g = labels_by_id[slot_to_id[s]];
        if (g == 0) continue;
        if (head) head[s] = first[g];
        if (tail) tail[s] = s == first[g] ? last[g] : (int32_t)s;
    }
    if (len) {
        for (int64_t i = 0; i < n; ++i) len[i] = 0;
        for (int64_t g = 1; g <= ngroups && g < n; ++g) len[g] = cnt[g];
    }
    return NBX_OK;
}
