/*
 * hfx_graph.h - whole pipelines from the reference's graph JSON, run by the library's
 * holoflow runtime (the node kinds of include/hfx_nodes.h, HIP streams, section threads).
 *
 * Graph JSON (src/holoflow/src/core/graph_spec.cc:66-170):
 *   {"nodes": {"<name>": {"type": "<kind>", "params": {...}}, ...},
 *    "edges": [{"from": "<name>", "to": "<name>", "out": <int>, "in": <int>}, ...]}
 * Compilation follows holoflow::runtime::Compiler (compiler.cc:288-1124) and execution the
 * Scheduler (graph_exec.cc:320-433).  Errors: HFX_E_INVALID for a malformed graph or
 * settings (messages as the reference's "GraphSpec JSON: ..." / CompilerException texts).
 */
#ifndef HFX_GRAPH_H_
#define HFX_GRAPH_H_

#include <stdint.h>

#include "hfx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfx_graph hfx_graph;

/* Compiler::compile (compiler.cc:288-356).  `prev` (may be NULL) is consumed, also on
 * failure: its threads are stopped and its buffers, streams and tasks are reused where the
 * new graph allows (tasks by name and kind through factory.update). */
hfx_status hfx_graph_compile(const char *graph_json, hfx_graph *prev, hfx_graph **out);
hfx_status hfx_graph_destroy(hfx_graph *g);
/* JSON: per node its kind, section and output descriptors / storage ids; the sections;
 * device and host bytes allocated by the compiler. */
hfx_status hfx_graph_describe(const hfx_graph *g, char *buf, int64_t cap);
/* Runs every section whose inputs are ready once, on the calling thread, in dependency
 * order (deterministic driving; not while the section threads run). */
hfx_status hfx_graph_step(hfx_graph *g, int32_t *sections_run);
/* Scheduler::start / request_stop + wait (graph_exec.cc:160-205): one thread per section;
 * stop returns the first error a section raised. */
hfx_status hfx_graph_start(hfx_graph *g);
hfx_status hfx_graph_stop(hfx_graph *g);
/* Per-node metrics JSON (average duration, runs/s, host and device bytes/s, samples). */
hfx_status hfx_graph_metrics(const hfx_graph *g, char *buf, int64_t cap);
/* Storage pointer and descriptor JSON of output `out_idx` of node `node` (host or device
 * memory as the descriptor's mem_loc says; null for an owned output between runs). */
hfx_status hfx_graph_output(const hfx_graph *g, const char *node, int32_t out_idx, void **data, char *desc_json,
                            int64_t cap);
/* core::to_dot of a graph JSON (graph_spec.cc:172-274). */
hfx_status hfx_graph_dot(const char *graph_json, char *buf, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* HFX_GRAPH_H_ */
