// Request loop of serve.hpp.
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <map>
#include <numeric>
#include <stdexcept>
#include <thread>

#include "../../include/gagp_hip.h"

namespace gagp_amd {

Server::Server(std::size_t budget_bytes, int poll_rounds) : budget_(budget_bytes), poll_rounds_(std::max(1, poll_rounds)) {}

int Server::submit(const SearchSpec& spec, double arrival_s) {
  requests_.push_back(Request{spec, arrival_s});
  ServeOutcome o;
  o.arrival_s = arrival_s;
  outcomes_.push_back(o);
  return int(requests_.size() - 1);
}

const ServeOutcome& Server::outcome(int ticket) const { return outcomes_.at(std::size_t(ticket)); }

void Server::run() {
  using Clock = std::chrono::steady_clock;
  const std::size_t n = requests_.size();
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return requests_[a].arrival_s < requests_[b].arrival_s; });
  Batch batch;
  // The round table changes at every admission and retirement; measuring
  // launch shapes for each new table would cost more than it returns.
  batch.set_autotune(0);
  std::map<int, int> ticket_of;  // batch index -> ticket of a live search
  std::map<int, std::size_t> bytes_of;
  std::size_t used = 0, next = 0, finished = 0;
  int pending = -1;              // batch index of the head request, parsed and planned
  std::vector<int> admitting;    // reserved in this pass, initialized together
  std::size_t pending_bytes = 0;
  const auto t0 = Clock::now();
  const auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  const auto fail = [&](ServeOutcome& o, int status, const std::string& what) {
    o.status = status;
    o.message = what;
    o.done_s = now();
    ++finished;
  };
  // Runs `f` for request `o`; any exception fails that request alone.
  const auto guarded = [&](ServeOutcome& o, auto&& f) {
    try {
      f();
      return true;
    } catch (const Unsupported& e) {
      fail(o, GAGP_ERR_UNSUPPORTED, e.what());
    } catch (const std::invalid_argument& e) {
      fail(o, GAGP_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
      fail(o, GAGP_ERR_DEVICE, e.what());
    }
    return false;
  };
  while (finished < n) {
    // Admission in arrival order while the head request fits (head-of-line:
    // a request that does not fit waits for memory, later ones wait behind
    // it).
    while (next < n && requests_[std::size_t(order[next])].arrival_s <= now()) {
      const int ticket = order[next];
      ServeOutcome& o = outcomes_[std::size_t(ticket)];
      if (pending < 0) {
        const bool planned = guarded(o, [&] {
          const int index = batch.add(requests_[std::size_t(ticket)].spec);
          pending = index;
          pending_bytes = batch.planned_footprint(index);
          if (pending_bytes > budget_)
            throw std::invalid_argument("search needs " + std::to_string(pending_bytes) + " device bytes, above the serving budget of " +
                                        std::to_string(budget_));
        });
        if (!planned) {
          if (pending >= 0) batch.drop(pending);
          pending = -1;
          ++next;
          continue;
        }
      }
      if (used + pending_bytes > budget_) break;
      if (max_live_ > 0 && int(ticket_of.size()) >= max_live_) break;
      // Reserved now, initialized below with every other request admitted
      // in this pass.
      ticket_of[pending] = ticket;
      bytes_of[pending] = pending_bytes;
      used += pending_bytes;
      admitting.push_back(pending);
      pending = -1;
      ++next;
    }
    if (!admitting.empty()) {
      std::vector<std::pair<int, std::string>> failed;
      try {
        failed = batch.admit_many(admitting);
      } catch (const std::exception& e) {
        for (int index : admitting) failed.push_back({index, e.what()});
      }
      for (const auto& f : failed) {
        fail(outcomes_[std::size_t(ticket_of.at(f.first))], GAGP_ERR_DEVICE, f.second);
        used -= bytes_of.at(f.first);
        ticket_of.erase(f.first);
        bytes_of.erase(f.first);
      }
      const double at = now();
      for (int index : admitting)
        if (ticket_of.count(index)) outcomes_[std::size_t(ticket_of.at(index))].admit_s = at;
      admitting.clear();
      peak_bytes_ = std::max(peak_bytes_, used);
      peak_live_ = std::max(peak_live_, int(ticket_of.size()));
    }
    if (ticket_of.empty()) {
      if (next < n) {
        const double wait = requests_[std::size_t(order[next])].arrival_s - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
      }
      continue;
    }
    batch.step(poll_rounds_);
    for (const Batch::Status& st : batch.status()) {
      if (!st.done) continue;
      ServeOutcome& o = outcomes_[std::size_t(ticket_of.at(st.index))];
      if (st.error) {
        o.status = GAGP_ERR_DEVICE;
        o.message = "device search error " + std::to_string(st.error);
      } else {
        try {
          const SearchResult r = batch.result(st.index, false);
          o.result_json = json::canonical(r.result);
          o.generations = int(r.result.at("generations").num);
        } catch (const std::exception& e) {
          o.status = GAGP_ERR_DEVICE;
          o.message = e.what();
        }
      }
      o.done_s = now();
      used -= bytes_of.at(st.index);
      batch.retire(st.index);
      ticket_of.erase(st.index);
      bytes_of.erase(st.index);
      ++finished;
    }
  }
}

}  // namespace gagp_amd
