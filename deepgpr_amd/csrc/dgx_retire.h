// Retire list: memory that queued GPU work may still read is released only
// once a completion marker recorded behind that work has completed.
//
// The batch runtime frees a finished instance's buffers and replaced
// descriptor tables through this list. Freeing them stream-ordered on the
// batch stream alone is not enough: buffers allocated later on another stream
// could then receive memory that steps queued on the batch stream still read.
//
// The list is a template over the marker operations so that a host-only test
// can drive it with fake events (tests/test_retire_list.py); the library
// instantiates it with HIP events (HipRetireOps in dgx_api.hip). Ops provides
//   using Event = ...; using Stream = ...;
//   bool create(Event*);          // a new marker
//   bool record(Event, Stream);   // the marker completes after the work queued on the stream so far
//   bool done(Event);             // the marker has completed
//   void destroy(Event);
//   bool sync(Stream);            // wait for all work queued on the stream
// Item is a movable owner whose destructor releases the memory.
#pragma once

#include <cstddef>
#include <utility>
#include <vector>

namespace dgx {

template <class Ops, class Item>
class RetireList {
  public:
    explicit RetireList(Ops ops = Ops()) : ops_(std::move(ops)) {}
    RetireList(const RetireList&) = delete;
    RetireList& operator=(const RetireList&) = delete;
    ~RetireList()
    {
        // the owner drains the stream before destroying the list
        for (Entry& e : entries_)
            if (e.has_event) ops_.destroy(e.event);
    }

    // Retire `item` behind a marker recorded on `stream` now. Returns 0, or -1
    // when no marker could be recorded: then the stream is drained first and
    // the item released at once, or, if even that fails, the item is kept
    // until the list is destroyed (never released while work may read it).
    int retire(typename Ops::Stream stream, Item item)
    {
        Entry e{std::move(item), typename Ops::Event(), false};
        if (ops_.create(&e.event)) {
            e.has_event = true;
            if (ops_.record(e.event, stream)) {
                entries_.push_back(std::move(e));
                return 0;
            }
        }
        if (ops_.sync(stream)) {
            if (e.has_event) ops_.destroy(e.event);
            return -1;  // e.item is released on return: nothing queued reads it any more
        }
        if (e.has_event) {
            ops_.destroy(e.event);
            e.has_event = false;
        }
        kept_.push_back(std::move(e.item));
        return -1;
    }

    // Release every item whose marker has completed; returns how many.
    size_t sweep()
    {
        size_t released = 0;
        for (size_t q = 0; q < entries_.size();) {
            if (!ops_.done(entries_[q].event)) {
                ++q;
                continue;
            }
            ops_.destroy(entries_[q].event);
            entries_.erase(entries_.begin() + (long)q);  // releases the item
            ++released;
        }
        return released;
    }

    size_t pending() const { return entries_.size() + kept_.size(); }

  private:
    struct Entry {
        Item item;
        typename Ops::Event event;
        bool has_event;
    };
    Ops ops_;
    std::vector<Entry> entries_;
    std::vector<Item> kept_;
};

}  // namespace dgx
