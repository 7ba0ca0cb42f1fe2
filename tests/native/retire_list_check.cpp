// Host-only driver of dgx::RetireList (deepgpr_amd/csrc/dgx_retire.h) with
// fake completion markers: checks that retired memory is released only after
// its marker has completed, and that a marker that cannot be created or
// recorded makes the list drain the stream (or keep the memory) first.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>
#include <vector>

#include "dgx_retire.h"

namespace {

std::vector<std::string> g_log;  // events in order: "release N", "sync", ...

struct FakeOps {
    using Event = int;
    using Stream = int;
    std::set<int>* completed;
    int* next;
    bool fail_create = false, fail_record = false, fail_sync = false;
    bool create(Event* e)
    {
        if (fail_create) return false;
        *e = (*next)++;
        return true;
    }
    bool record(Event e, Stream) { return !fail_record && e >= 0; }
    bool done(Event e) { return completed->count(e) != 0; }
    void destroy(Event e) { g_log.push_back("destroy " + std::to_string(e)); }
    bool sync(Stream)
    {
        g_log.push_back("sync");
        return !fail_sync;
    }
};

struct Mem {  // stands in for device memory; its destructor is the release
    int id = -1;
    Mem() = default;
    explicit Mem(int i) : id(i) {}
    Mem(Mem&& o) noexcept : id(o.id) { o.id = -1; }
    Mem& operator=(Mem&& o) noexcept
    {
        if (this != &o) {
            if (id >= 0) g_log.push_back("release " + std::to_string(id));
            id = o.id;
            o.id = -1;
        }
        return *this;
    }
    ~Mem()
    {
        if (id >= 0) g_log.push_back("release " + std::to_string(id));
    }
};

int g_fail = 0;
void expect(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        g_fail = 1;
    }
}
bool logged(const std::string& s)
{
    for (auto& l : g_log)
        if (l == s) return true;
    return false;
}
size_t index_of(const std::string& s)
{
    for (size_t i = 0; i < g_log.size(); ++i)
        if (g_log[i] == s) return i;
    return (size_t)-1;
}

}  // namespace

int main()
{
    std::set<int> completed;
    int next = 0;
    {
        FakeOps ops{&completed, &next};
        dgx::RetireList<FakeOps, Mem> list(ops);
        expect(list.retire(0, Mem(1)) == 0, "retire with a marker succeeds");
        expect(list.retire(0, Mem(2)) == 0, "second retire succeeds");
        expect(!logged("release 1") && !logged("release 2"), "nothing released at retire time");
        expect(list.sweep() == 0, "sweep before completion releases nothing");
        expect(!logged("release 1") && !logged("release 2"), "still held while markers pend");
        completed.insert(1);  // the second marker completes first
        expect(list.sweep() == 1, "one marker completed, one release");
        expect(logged("release 2") && !logged("release 1"), "only the completed item is released");
        completed.insert(0);
        expect(list.sweep() == 1, "the other completes");
        expect(logged("release 1"), "released after completion");
        expect(list.pending() == 0, "empty");
    }
    g_log.clear();
    {
        FakeOps ops{&completed, &next};
        ops.fail_record = true;
        dgx::RetireList<FakeOps, Mem> list(ops);
        expect(list.retire(0, Mem(3)) == -1, "a failed record is reported");
        const size_t s = index_of("sync"), r = index_of("release 3");
        expect(s != (size_t)-1 && r != (size_t)-1 && s < r, "the stream is drained before the release");
        expect(list.pending() == 0, "nothing left after a drained release");
    }
    g_log.clear();
    {
        FakeOps ops{&completed, &next};
        ops.fail_create = true;
        ops.fail_sync = true;
        {
            dgx::RetireList<FakeOps, Mem> list(ops);
            expect(list.retire(0, Mem(4)) == -1, "a failed create is reported");
            expect(!logged("release 4"), "kept while the stream cannot be drained");
            expect(list.pending() == 1, "kept in the list");
            expect(list.sweep() == 0, "a kept item has no marker to complete");
            expect(!logged("release 4"), "still kept after a sweep");
        }
        expect(logged("release 4"), "released with the list (after the owner drained the stream)");
    }
    if (!g_fail) std::printf("retire list: all checks passed\n");
    return g_fail;
}
