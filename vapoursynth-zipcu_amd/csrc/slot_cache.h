// Lease table: the bookkeeping of keyed buffers shared by concurrent frame requests. It gives the
// guarantees src/framecache.zig gives its device frame cache: a request gets all of its keys at
// once or waits holding none; a key that is resident is shared, one that is not gets the least
// recently used free buffer and exactly one request fills it; a claim that is never filled is
// forgotten, so the next request for that key fills it again.
//
// Design: a request waits until none of its keys is being filled by another request and enough
// unpinned buffers exist for its missing keys, then pins and claims them all under one lock. The
// returned Lease unpins its buffers when it ends and forgets every buffer it was to fill but did
// not. No HIP here: buffer i of the table is whatever the owner keeps at index i.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <unordered_map>
#include <vector>

class LeaseTable {
    enum class State : uint8_t { Empty, Filling, Resident };
    struct Meta {
        int64_t key = -1;
        State state = State::Empty;
        int pins = 0;
        uint64_t used = 0;  // tick of the last acquisition (0: never, or forgotten)
    };

  public:
    class Lease {
      public:
        Lease() = default;
        Lease(const Lease &) = delete;
        Lease &operator=(const Lease &) = delete;
        Lease(Lease &&o) noexcept { take(o); }
        Lease &operator=(Lease &&o) noexcept {
            if (this != &o) {
                end();
                take(o);
            }
            return *this;
        }
        ~Lease() { end(); }

        bool held() const { return table_ != nullptr; }
        // Buffer index of request key i.
        int slot(int i) const { return slot_[i]; }
        // True when this lease must fill the buffer of request key i (the first request position
        // of a key that was not resident).
        bool must_fill(int i) const { return fill_[i] != 0; }
        // The buffer of request key i now holds its key's data: other requests may use it.
        void filled(int i) {
            {
                std::lock_guard<std::mutex> lk(table_->mu_);
                table_->meta_[slot_[i]].state = State::Resident;
            }
            fill_[i] = 0;
            table_->cv_.notify_all();
        }
        // Unpins every buffer; buffers still to be filled become empty again. Idempotent.
        void end() {
            if (!table_) return;
            {
                std::lock_guard<std::mutex> lk(table_->mu_);
                for (size_t i = 0; i < slot_.size(); ++i)
                    if (fill_[i]) table_->forget(slot_[i]);
                for (int id : pinned_) table_->meta_[id].pins -= 1;
            }
            table_->cv_.notify_all();
            table_ = nullptr;
            slot_.clear();
            fill_.clear();
            pinned_.clear();
        }

      private:
        friend class LeaseTable;
        void take(Lease &o) {
            table_ = o.table_;
            slot_ = std::move(o.slot_);
            fill_ = std::move(o.fill_);
            pinned_ = std::move(o.pinned_);
            o.table_ = nullptr;
        }
        LeaseTable *table_ = nullptr;
        std::vector<int> slot_;    // per request position
        std::vector<char> fill_;   // per request position
        std::vector<int> pinned_;  // distinct buffers, one pin each
    };

    // Forgets every key; n buffers. Only while no lease is held.
    void reset(size_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        meta_.assign(n, Meta{});
        index_.clear();
        tick_ = 0;
    }

    size_t size() const { return meta_.size(); }

    // Blocks until every key has a pinned buffer. A lease that is not held() means the request
    // has more distinct keys than the table has buffers (it could never be granted).
    Lease lease(const int64_t *keys, int n) {
        std::vector<int64_t> distinct;
        std::vector<int> pos(n);
        for (int i = 0; i < n; ++i) {
            const auto it = std::find(distinct.begin(), distinct.end(), keys[i]);
            pos[i] = int(it - distinct.begin());
            if (it == distinct.end()) distinct.push_back(keys[i]);
        }
        Lease L;
        std::unique_lock<std::mutex> lk(mu_);
        if (distinct.size() > meta_.size()) return L;
        cv_.wait(lk, [&] { return grantable(distinct); });
        // pin the resident keys first, so that none of them is taken as a victim below
        std::vector<int> dslot(distinct.size(), -1);
        std::vector<char> dfill(distinct.size(), 0);
        for (size_t d = 0; d < distinct.size(); ++d) {
            const auto it = index_.find(distinct[d]);
            if (it == index_.end()) continue;
            dslot[d] = it->second;
            meta_[it->second].pins += 1;
            meta_[it->second].used = ++tick_;
        }
        for (size_t d = 0; d < distinct.size(); ++d) {
            if (dslot[d] >= 0) continue;
            const int v = victim();  // exists: grantable() counted enough free buffers
            forget(v);
            meta_[v].key = distinct[d];
            meta_[v].state = State::Filling;
            meta_[v].pins = 1;
            meta_[v].used = ++tick_;
            index_[distinct[d]] = v;
            dslot[d] = v;
            dfill[d] = 1;
        }
        L.table_ = this;
        L.pinned_ = dslot;
        L.slot_.resize(n);
        L.fill_.assign(n, 0);
        for (int i = 0; i < n; ++i) {
            L.slot_[i] = dslot[pos[i]];
            if (dfill[pos[i]]) {
                L.fill_[i] = 1;
                dfill[pos[i]] = 0;  // later positions of the same key share the first one's duty
            }
        }
        return L;
    }

  private:
    // No key of the request is being filled by another request, and the unpinned buffers that
    // hold none of its keys can take every missing key.
    bool grantable(const std::vector<int64_t> &distinct) const {
        int missing = 0;
        for (int64_t k : distinct) {
            const auto it = index_.find(k);
            if (it == index_.end()) ++missing;
            else if (meta_[it->second].state == State::Filling) return false;
        }
        int free_bufs = 0;
        for (const Meta &m : meta_) {
            if (m.pins) continue;
            if (m.state != State::Empty && std::find(distinct.begin(), distinct.end(), m.key) != distinct.end())
                continue;
            ++free_bufs;
        }
        return free_bufs >= missing;
    }

    // Least recently used unpinned buffer (empty ones have the oldest tick, 0).
    int victim() const {
        int v = -1;
        for (size_t i = 0; i < meta_.size(); ++i) {
            if (meta_[i].pins) continue;
            if (v < 0 || meta_[i].used < meta_[v].used) v = int(i);
        }
        return v;
    }

    void forget(int id) {
        Meta &m = meta_[id];
        if (m.key >= 0) {
            const auto it = index_.find(m.key);
            if (it != index_.end() && it->second == id) index_.erase(it);
        }
        m.key = -1;
        m.state = State::Empty;
        m.used = 0;
    }

    std::vector<Meta> meta_;
    std::unordered_map<int64_t, int> index_;
    std::mutex mu_;
    std::condition_variable cv_;
    uint64_t tick_ = 0;
};
