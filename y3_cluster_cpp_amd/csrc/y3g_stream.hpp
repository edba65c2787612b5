// y3g_stream.hpp -- the runtime's continuous-batching session, internal to
// liby3g (the request server of y3g_server.cpp drives it; the functions are
// not exported from the library).
//
// A session owns one runtime's tables (table_slots slots) and region slab,
// a ring of ring_items entries in coherent pinned host memory, and at most
// one running k_cubature_stream launch.  The host publishes table builds
// and integrals into the ring; the launch claims them in order as
// workgroups free up, so work published while the launch runs joins it.
// An idle launch ends by itself (idle_ms); stream_kick starts another when
// published work is left.
#pragma once

#include <stddef.h>

#include "y3g.h"

struct y3g_stream;

namespace y3g {

#define Y3G_HIDDEN __attribute__((visibility("hidden")))

struct StreamConfig {
  int ring_items;        // ring entries (integrals and table builds in flight)
  int table_slots;       // table sets kept on the device
  double idle_ms;        // an idle launch ends after this long
  double table_wait_ms;  // bound on an integral's wait for its tables
  y3g_options options;   // cubature options (precision FP32SPLIT or FP64)
  y3g_table_spec spec;
};

// Device bytes a session on rt would take: the region slab of its resident
// workgroups, and one table slot.
Y3G_HIDDEN int stream_footprint(y3g_runtime* rt, const StreamConfig& c, size_t* slab_bytes,
                                size_t* slot_bytes, int* workgroups);
Y3G_HIDDEN int stream_open(y3g_runtime* rt, const StreamConfig& c, y3g_stream** out);
// Free ring entries.
Y3G_HIDDEN size_t stream_room(const y3g_stream* s);
// Table builds of a slot published and not yet finished.
Y3G_HIDDEN int stream_tables_pending(const y3g_stream* s, int slot);
// Publish a build of slot's tables for p; *gen = the generation it makes.
Y3G_HIDDEN int stream_tables(y3g_stream* s, int slot, const y3g_params& p, unsigned* gen);
// Publish n integrals on slot (generation gen); *first = the first entry.
Y3G_HIDDEN int stream_push(y3g_stream* s, int slot, unsigned gen, const y3g_integral* igs,
                           size_t n, unsigned long long* first);
// Report the integrals finished since the last call, f(ctx, entry, result)
// each; returns how many.
Y3G_HIDDEN size_t stream_collect(y3g_stream* s,
                                 void (*f)(void*, unsigned long long, const y3g_result&),
                                 void* ctx);
// Start a launch when published entries are unclaimed and none runs.
Y3G_HIDDEN int stream_kick(y3g_stream* s);
// Launches started so far.
Y3G_HIDDEN long long stream_launches(const y3g_stream* s);
// End the launch (every published entry is claimed first) and free the
// session; *nevals = integrand evaluations of all launches.
Y3G_HIDDEN int stream_close(y3g_stream* s, long long* nevals);

}  // namespace y3g
