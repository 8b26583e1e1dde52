// Serving loop: simulation requests arrive at any time, are admitted to the
// device while their footprint fits an HBM budget, and run together in
// rounds (continuous batching).  Each request is one output::run_with
// (crates/cpm-core/src/output.rs:434-529).
#pragma once

#include <cstdint>
#include <string>

#include "device.hpp"
#include "host.hpp"

namespace cpm {

struct Server;

struct ServeOptions {
    uint64_t hbm_budget = 0;   // bytes of device memory requests may hold
    uint32_t max_active = 0;   // requests on the device at once
    uint64_t quantum = 0;      // attempts per sublattice-mode request per round
    double round_s = 0;        // device time of an exact-replay chain per round
};

struct RequestInfo {
    int state = 0;  // CPM_REQ_*
    double submit_s = 0, admit_s = 0, done_s = 0;
    uint64_t footprint = 0, mcs_done = 0, mcs_total = 0;
    std::string error;
};

struct ServerStats {
    uint64_t submitted = 0, completed = 0, failed = 0, queued = 0, active = 0;
    uint64_t hbm_budget = 0, hbm_in_use = 0, rounds = 0, attempts = 0;
    double busy_s = 0, device_s = 0, uptime_s = 0;
};

Server* serve_create(Engine* e, const ServeOptions& o);
// Stops the worker after its current round; requests not finished are
// dropped.
void serve_destroy(Server* s);
// Takes ownership of an initialised model; throws Error when the request
// can never run (footprint above the budget, unsupported mode).
uint64_t serve_submit(Server* s, Model* m, bool include_lattice, int mode);
// Waits up to timeout_s (0: no wait) for the request to finish.  Returns
// true and moves the result into `out` when it is done (the request is then
// forgotten); throws Error for unknown ids and for failed requests (which
// are forgotten too).
bool serve_poll(Server* s, uint64_t id, double timeout_s, RequestInfo& info, Result* out);
ServerStats serve_stats(Server* s);

}  // namespace cpm
