// In-flight bookkeeping of a pipelined item processor (the Service contract of
// service/Service.h:24-70, worker loop server/Worker.tpp:13-34): items are
// issued into slots and retired in issue order as their work completes.
#pragma once

#include <deque>

namespace as {

enum class SlotStatus { Done, Pending, Failed };

// Retires the slots at the front of `inflight` (slot indices in issue order)
// whose work has ended.  query(slot, block) reports a slot's state, waiting for
// it when block is true (asked for the oldest slot only); done(slot) marks the
// slot's item processed and frees the slot; drop(slot) frees a failed slot
// without processing its item (the item-processor contract has no error
// channel: the client's wait() reports the request).  A failed slot never
// stays in flight, so a processor that returns "items in flight" while
// `inflight` is non-empty always drains.  Returns false if a slot failed.
template <class Query, class Done, class Drop>
bool retire_in_order(std::deque<int>& inflight, bool block, Query query, Done done, Drop drop) {
    bool ok = true;
    while (!inflight.empty()) {
        const int k = inflight.front();
        const SlotStatus st = query(k, block);
        if (st == SlotStatus::Pending) return ok;
        inflight.pop_front();
        if (st == SlotStatus::Done) {
            done(k);
        } else {
            drop(k);
            ok = false;
        }
        block = false;
    }
    return ok;
}

}  // namespace as
