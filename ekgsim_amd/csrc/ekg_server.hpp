// ekg_server.hpp -- continuous-batching request loop (ekg_server_* in ekg.h).
#pragma once

#include "ekg.h"
#include "ekg_engine.hpp"

namespace ekg {

struct Server;

Server* server_create(Ctx& ctx, const ekg_server_opts& opts);
void server_destroy(Server* s);
int server_add_heart(Server& s, Heart& h);
uint64_t server_submit(Server& s, int heart_id, const ekg_sim_desc& d);
bool server_poll(Server& s, uint64_t ticket);
void server_wait(Server& s, uint64_t ticket, double* ecg, size_t cap, double* latency_ms);
// Waits like server_wait and returns the result in place (n doubles) instead
// of copying it; the ticket stays until server_release.
const double* server_wait_view(Server& s, uint64_t ticket, size_t* n, double* latency_ms);
void server_release(Server& s, uint64_t ticket);
void server_stats(Server& s, uint64_t* completed, uint64_t* groups, int* peak_inflight,
                  size_t* peak_bytes);

}  // namespace ekg
