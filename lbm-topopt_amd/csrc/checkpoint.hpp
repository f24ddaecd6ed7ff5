// Online level-based checkpoint placement with recomputation (Wang, Moin & Iaccarino 2009),
// the schedule of the reference's DynamicCheckpointing + WangCheckpointing
// (src/topOpt/dynamicCheckpointing.cc:9-54, src/topOpt/WangCheckpointing.cc:36-126), and its
// StaticCheckpointing variant for a known horizon (src/topOpt/staticCheckpointing.cc:3-55).
// A third placement, binomial checkpointing for a known horizon (Griewank 1992; Griewank &
// Walther, ACM TOMS 26 (2000) 19-45, "revolve"), has no counterpart in the reference: it
// delivers the same states with the fewest forward steps any placement of the slots allows.
//
// Storage is abstract: the schedule hands out slot ids 0..nslots-1 and the engine maps each
// slot to an HBM buffer. Placing a state in a slot is done by exchanging buffer handles
// (no device copies), so the schedule only has to say which step goes to which slot.
#pragma once
#include <algorithm>
#include <climits>
#include <functional>
#include <list>
#include <stdexcept>
#include <vector>

namespace lbmx {

class WangSchedule {
 public:
  struct Entry {
    int step;
    int level;
    int slot;
  };
  // A request for the state at `index`: restart from `from_step` held in `from_slot`,
  // advance, and after producing each step in `stores` put the state into that slot.
  // The state at `index` itself is returned to the caller (not stored).
  struct Plan {
    int from_step = 0;
    int from_slot = -1;
    int index = 0;
    std::vector<std::pair<int, int>> stores;  // (step, slot), increasing step
    int advances() const { return index - from_step; }
    bool stored() const { return from_step == index; }  // no recomputation needed
  };

  explicit WangSchedule(int nslots) : nslots_(nslots) {
    if (nslots < 2) throw std::invalid_argument("dynamic checkpointing needs at least 2 slots");
    reset();
  }

  int nslots() const { return nslots_; }

  void reset() {
    list_.clear();
    free_.clear();
    for (int s = nslots_ - 1; s >= 0; --s) free_.push_back(s);
    have_reader_ = false;
    next_planned_ = 0;
    next_binomial_ = 0;
  }

  // StaticCheckpointing: the forward pass keeps exactly the checkpoints that the level-based
  // placement holds after planning the whole horizon 0..last at once (computeNewCheckpoints
  // from the initial state), with their levels; the reverse pass is the same as for the
  // dynamic schedule.
  void plan_static(int last) {
    reset();
    list_.push_front({0, INT_MAX, take()});
    place(0, last);
    planned_.clear();
    for (auto it = list_.rbegin(); it != list_.rend(); ++it) planned_.push_back({it->step, it->level, -1});
    reset();
    static_ = true;
  }
  bool is_static() const { return static_; }

  // Minimal number of forward steps that reverse a chain of n states (the adjoint of the
  // chain's last step needs state n-1, the first needs state 0) with c checkpoints, one of them
  // holding state 0, counting the first pass through the chain: r n - C(c+r, c+1) with r the
  // least integer such that C(c+r, c) >= n (Griewank & Walther 2000, Proposition 1).
  static long long binomial_cost(long long n, int c) {
    if (n <= 1) return 0;
    if (c < 1) throw std::invalid_argument("binomial checkpointing needs a checkpoint");
    long long r = 0, b = 1;  // b = C(c + r, c)
    while (b < n) {
      ++r;
      b = b * (c + r) / r;
    }
    return r * n - b * r / (c + 1);  // C(c+r, c+1) = C(c+r, c) r / (c+1)
  }

  // Binomial placement for a known horizon of n states 0..n-1. The forward pass keeps the
  // checkpoints of the first descent; every later request is planned from the schedule built
  // here. With x_a held and s free slots, a chain of n states splits at the k that minimises
  // k + cost(n-k, s) + cost(k, s+1): advance k steps and keep state a+k, reverse the tail with
  // one slot fewer, free that slot, then reverse the head. The cost is convex in k, so a binary
  // search on its forward difference finds the split.
  void plan_binomial(int n) {
    if (n < 1) throw std::invalid_argument("binomial checkpointing needs at least one state");
    reset();
    static_ = false;
    binomial_ = true;
    fwd_slot_.assign((size_t)n, -1);
    bplans_.clear();
    std::vector<int> pool;
    for (int s = nslots_ - 1; s >= 0; --s) pool.push_back(s);
    auto grab = [&]() {
      int s = pool.back();
      pool.pop_back();
      return s;
    };
    Plan cur;
    auto restore = [&](int a, int sa) {
      cur = Plan();
      cur.from_step = a;
      cur.from_slot = sa;
    };
    auto deliver = [&](int i, int release) {
      cur.index = i;
      bplans_.push_back({cur, release});
    };
    auto split = [](long long len, int s) {
      auto f = [&](long long k) { return k + binomial_cost(len - k, s) + binomial_cost(k, s + 1); };
      long long lo = 1, hi = len - 1;  // smallest k in [1, len-1] with f(k+1) >= f(k)
      while (lo < hi) {
        long long mid = lo + (hi - lo) / 2;
        if (f(mid + 1) >= f(mid)) hi = mid;
        else lo = mid + 1;
      }
      return (int)lo;
    };
    // reverse(a, sa, len, s, continuing): x_a is in slot sa, s slots are free; continuing means
    // the chain being built already passes through x_a (no restore needed)
    std::function<void(int, int, int, int, bool)> reverse = [&](int a, int sa, int len, int s, bool continuing) {
      for (;;) {
        if (len == 1) {
          if (!continuing) restore(a, sa);
          deliver(a, sa);
          return;
        }
        if (s == 0) {
          for (int i = a + len - 1; i > a; --i) {
            if (!continuing) restore(a, sa);
            continuing = false;
            deliver(i, -1);
          }
          restore(a, sa);
          deliver(a, sa);
          return;
        }
        const int k = split(len, s);
        const int sk = grab();
        if (!continuing) restore(a, sa);
        cur.stores.push_back({a + k, sk});
        reverse(a + k, sk, len - k, s - 1, true);  // frees sk when it delivers a+k
        pool.push_back(sk);
        len = k;  // the head, with the same free slots
        continuing = false;
      }
    };
    const int s0 = grab();
    fwd_slot_[0] = s0;
    restore(0, s0);
    reverse(0, s0, n, nslots_ - 1, true);
    // The first plan is the forward pass itself: its stores are the forward checkpoints, and
    // the state it delivers (n-1) is the forward pass's last input, kept by the caller, unless
    // the descent stored it.
    BinomialPlan& first = bplans_.front();
    for (auto& st : first.plan.stores) fwd_slot_[(size_t)st.first] = st.second;
    const bool kept = !first.plan.stores.empty() && first.plan.stores.back().first == n - 1;
    first.plan.from_slot = kept ? first.plan.stores.back().second : (n == 1 ? s0 : -1);
    first.plan.from_step = n - 1;
    first.plan.stores.clear();
  }
  bool is_binomial() const { return binomial_; }

  // DynamicCheckpointing::makeCheckpoint: state `index` always becomes the newest checkpoint.
  // Returns its slot. The slot's previous content (if any) is dropped. A static schedule keeps
  // only its planned steps (StaticCheckpointing::makeCheckpoint) and returns -1 for the others.
  int make_checkpoint(int index) {
    if (binomial_) {
      if (index < 0 || (size_t)index >= fwd_slot_.size() || fwd_slot_[(size_t)index] < 0) return -1;
      const int s = fwd_slot_[(size_t)index];
      claim(s);
      return s;
    }
    if (static_) {
      if (next_planned_ >= planned_.size() || planned_[next_planned_].step != index) return -1;
      int s = take();
      list_.push_front({index, planned_[next_planned_].level, s});
      ++next_planned_;
      return s;
    }
    if (list_.empty()) {
      int s = take();
      list_.push_front({index, INT_MAX, s});
      reader_ = list_.begin();
      have_reader_ = true;
      return s;
    }
    if (!free_.empty()) {
      int s = take();
      list_.push_front({index, 0, s});
      reader_ = list_.begin();
      have_reader_ = true;
      return s;
    }
    auto movable = std::next(reader_);
    if (movable == list_.end()) throw std::logic_error("checkpoint schedule exhausted");
    if (movable->level < reader_->level) {
      int s = movable->slot;
      list_.push_front({index, 0, s});
      list_.erase(movable);
      return s;
    }
    reader_ = list_.begin();
    int lvl = reader_->level + 1;
    int s = reader_->slot;
    list_.push_front({index, lvl, s});
    list_.erase(reader_);
    reader_ = list_.begin();
    return s;
  }

  // WangCheckpointing::getForwardSolution: plans the delivery of state `index`, placing new
  // checkpoints between the newest stored state and `index` (computeNewCheckpoints). The
  // checkpoint of `index` is consumed; call release() once the state has been used.
  Plan get_forward_solution(int index) {
    if (binomial_) {
      if (next_binomial_ >= bplans_.size() || bplans_[next_binomial_].plan.index != index)
        throw std::logic_error("states must be requested in decreasing order");
      const BinomialPlan& bp = bplans_[next_binomial_++];
      for (auto& st : bp.plan.stores) claim(st.second);
      consumed_slot_ = bp.release;
      return bp.plan;
    }
    Plan p;
    p.index = index;
    if (list_.empty()) throw std::logic_error("no checkpoint stored");
    const Entry start = list_.front();
    if (start.step > index) throw std::logic_error("states must be requested in decreasing order");
    p.from_step = start.step;
    p.from_slot = start.slot;
    if (start.step == index) {
      consumed_slot_ = start.slot;
      list_.pop_front();
      return p;
    }
    place(start.step, index);
    // new entries are those in front of the (former) starting checkpoint, all with step >
    // start.step; the newest one is `index`.
    for (auto& e : list_) {
      if (e.step <= start.step) break;
      if (e.step < index) p.stores.push_back({e.step, e.slot});
    }
    std::reverse(p.stores.begin(), p.stores.end());
    if (list_.front().step != index) throw std::logic_error("schedule lost the requested state");
    consumed_slot_ = list_.front().slot;
    list_.pop_front();
    return p;
  }

  // free the slot consumed by the last get_forward_solution
  void release() {
    if (consumed_slot_ >= 0) free_.push_back(consumed_slot_);
    consumed_slot_ = -1;
  }

  const std::list<Entry>& entries() const { return list_; }
  int free_slots() const { return static_cast<int>(free_.size()); }

 private:
  int take() {
    int s = free_.back();
    free_.pop_back();
    return s;
  }
  // mark a slot named by the binomial plan as in use; it must be free
  void claim(int s) {
    auto it = std::find(free_.begin(), free_.end(), s);
    if (it == free_.end()) throw std::logic_error("binomial plan reuses a slot in use");
    free_.erase(it);
  }

  // computeNewCheckpoints (src/topOpt/WangCheckpointing.cc:83-126)
  void place(int from, int index) {
    auto reader = list_.begin();
    for (int ii = from; ii < index; ++ii) {
      if (!free_.empty()) {
        list_.push_front({ii + 1, 0, take()});
        reader = list_.begin();
        continue;
      }
      auto movable = std::next(reader);
      if (movable == list_.end()) throw std::logic_error("checkpoint schedule exhausted");
      if (movable->level < reader->level) {
        int s = movable->slot;
        list_.push_front({ii + 1, 0, s});
        list_.erase(movable);
      } else {
        reader = list_.begin();
        int lvl = reader->level + 1;
        int s = reader->slot;
        list_.push_front({ii + 1, lvl, s});
        list_.erase(reader);
        reader = list_.begin();
      }
    }
    // the dynamic placement's reader refers to forward-pass entries only
    have_reader_ = false;
  }

  struct BinomialPlan {
    Plan plan;
    int release;  // slot freed once the delivered state has been used, or -1
  };
  int nslots_;
  bool static_ = false;
  bool binomial_ = false;
  std::vector<int> fwd_slot_;         // binomial: slot of each forward checkpoint, or -1
  std::vector<BinomialPlan> bplans_;  // binomial: requests in order (decreasing index)
  size_t next_binomial_ = 0;
  std::vector<Entry> planned_;  // static schedule: forward checkpoints, increasing step
  size_t next_planned_ = 0;
  std::list<Entry> list_;  // newest first
  std::vector<int> free_;
  std::list<Entry>::iterator reader_;
  bool have_reader_ = false;
  int consumed_slot_ = -1;
};

}  // namespace lbmx
