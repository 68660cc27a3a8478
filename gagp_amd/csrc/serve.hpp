// Serving: a request loop with continuous batching over the search engine.
// Requests arrive over time; each is admitted into the running batch as soon
// as it has arrived and its device footprint -- computed from its population,
// batch, row length and case count before anything is allocated -- fits the
// budget beside the live searches (first come, first served), its population
// stays on the device for its whole life, and it retires -- publishing its
// result.json and returning its memory -- when it reaches its generation cap
// or, with early exit, the first generation boundary at which its best member
// has no wrong train case. A request that fails (invalid or unsupported
// documents, a footprint above the whole budget, a device allocation or
// initialization failure) fails alone: its ticket carries the status and
// message, and the loop serves the others.
#pragma once
#include <string>
#include <vector>

#include "engine.hpp"

namespace gagp_amd {

struct ServeOutcome {
  double arrival_s = 0, admit_s = -1, done_s = -1;  // seconds from the start of run()
  int generations = 0;                              // generations of the published result
  int status = 0;                                   // 0 ok, else a GAGP_ERR_* status
  std::string message, result_json;
};

class Server {
 public:
  // budget_bytes: device memory the live searches may hold together;
  // poll_rounds: rounds stepped between completion checks.
  Server(std::size_t budget_bytes, int poll_rounds);
  // At most `n` live searches at once (0: no limit besides the budget);
  // 1 serves one search at a time.
  void set_max_live(int n) { max_live_ = n < 0 ? 0 : n; }
  int submit(const SearchSpec& spec, double arrival_s);
  void run();
  const ServeOutcome& outcome(int ticket) const;
  std::size_t size() const { return requests_.size(); }
  std::size_t peak_bytes() const { return peak_bytes_; }
  int peak_live() const { return peak_live_; }

 private:
  struct Request { SearchSpec spec; double arrival_s; };
  std::vector<Request> requests_;
  std::vector<ServeOutcome> outcomes_;
  std::size_t budget_, peak_bytes_ = 0;
  int poll_rounds_, peak_live_ = 0, max_live_ = 0;
};

}  // namespace gagp_amd
