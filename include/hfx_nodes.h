/*
 * hfx_nodes.h - C ABI over the holoflow node layer (include/holoflow/core/...).
 *
 * Runs any registered node kind through its factory exactly as the reference's
 * single-node harness does (test/holonp/support/sync_task_runner.hh:26-53: infer ->
 * create on a stream -> execute -> update), with tensor descriptors and settings as JSON
 * text.  Node kinds and settings keys are the reference's (manager.cc:159-232 and each
 * op's to_json/from_json); the fused kind "DopplerMoments" is added.
 *
 * Descriptor JSON: {"shape": [...], "dtype": "U8"|"U16"|"F32"|"CF32", "mem_loc": "Device",
 *                   "strides": [bytes...], "offset": bytes}   (strides/offset optional)
 * Data pointers passed to execute/acquire/pop are storage base pointers; a descriptor's
 * byte offset is applied by the node, as TView::data() does.
 */
#ifndef HFX_NODES_H_
#define HFX_NODES_H_

#include <stdint.h>

#include "hfx.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HFX_NOT_READY 5 /* OpResult::NotReady of an asynchronous node */
#define HFX_EOF 6       /* OpResult::Eof */

typedef struct hfx_node hfx_node;

/* JSON array of the registered kinds (Registry::kinds, registry.hh:24-68). */
hfx_status hfx_node_kinds(char *buf, int64_t cap);

/* ISyncTaskFactory/IAsyncTaskFactory::infer (tasks.hh:260-275).  Writes
 * {"kind": "Sync"|"Async", "output_descs": [...], "in_place": [[in, out], ...],
 *  "owned_inputs": [...], "owned_outputs": [...]} to buf. */
hfx_status hfx_node_infer(const char *kind, const char *settings_json, const char *input_descs_json,
                          char *buf, int64_t cap);

/* factory.create(inputs, settings, {stream}) (tasks.hh:276-290, 311-325). */
hfx_status hfx_node_create(const char *kind, const char *settings_json, const char *input_descs_json,
                           hfx_stream_t stream, hfx_node **out);

/* factory.update(old, ...) (tasks.hh:292-305, 327-334); *reused = 1 when the node kept its
 * task (e.g. FresnelDiffraction with only z changed, fresnel_diffraction.cu:634-668). */
hfx_status hfx_node_update(hfx_node *node, const char *settings_json, const char *input_descs_json,
                           hfx_stream_t stream, int32_t *reused);

/* ISyncTask::execute (tasks.hh:182-190) with storage base pointers for the inferred inputs
 * and outputs (host or device memory, as each descriptor's mem_loc says); work is enqueued
 * on the node's stream.  For an output the node owns (infer's owned_outputs, e.g. a cached
 * Holofile source) outputs[i] receives the pointer of the view it hands out instead. */
hfx_status hfx_node_execute(hfx_node *node, void *const *inputs, int32_t n_inputs, void **outputs,
                            int32_t n_outputs);

/* IAsyncTask owned slots and queue operations (tasks.hh:136-151, 208-220):
 * acquire_input -> producer writes the slot -> push; pop -> consumer reads -> release.
 * push: extra_inputs are the inputs after the owned first one (e.g. SlidingAverage's host
 * U8 validity flag), as storage base pointers.  pop: outputs[i] receives the storage
 * pointer of each output the node owns; for an output it does not own (DualReaderBatchQueue's
 * host U8 validity flag) the caller passes the buffer the node writes.  HFX_NOT_READY when
 * nothing can be popped.  release_output(index) releases one owned output (of an async node,
 * or of a sync node that owns its outputs). */
hfx_status hfx_node_acquire_input(hfx_node *node, void **slot);
hfx_status hfx_node_push(hfx_node *node, void *const *extra_inputs, int32_t n_extra);
hfx_status hfx_node_pop(hfx_node *node, void **outputs, int32_t n_outputs);
hfx_status hfx_node_release_output(hfx_node *node, int32_t index);

hfx_status hfx_node_destroy(hfx_node *node);

#ifdef __cplusplus
}
#endif

#endif /* HFX_NODES_H_ */
