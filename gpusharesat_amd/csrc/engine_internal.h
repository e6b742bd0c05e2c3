// Library-internal access to an engine's device state, for modules built into
// the same shared object (the residency scorer).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "gss.h"
#include "kernels.h"

namespace gss {

// Compacts a tagged DB on the device keeping rows whose tag t has d_keep[t]
// != 0 (d_keep is device memory; tags >= ntags are kept).
int db_compact_device(gss_engine *e, const uint8_t *d_keep, int64_t ntags);
// Device tag array of every bucket, indexed by clause length (null if none).
int db_tag_arrays(gss_engine *e, std::vector<const uint32_t *> &tags);
// The chunked hit records of the last waited plain sweep, on the device:
// chunk c holds fills[c] records starting at hits + c * kHitChunk.  Refused
// once rows have moved since that sweep (compaction, bucket replacement,
// clear), since the records' (len, row) would no longer name the same rows.
int sweep_records(gss_engine *e, const gss_hit **hits, const uint32_t **fills,
                  uint32_t *chunks);
hipStream_t engine_stream(gss_engine *e);
// Serving (gss_server): attaches the request and patch rings (fine-grained
// pinned host memory the host publishes into, head = requests published) to
// the engine; the device-side admission cursor starts at request 0.
// Pinned record buffers per step slot (1 .. 4): a waited step's records then
// stay in place for that many waits of its slot (2n steps with two slots).
int engine_set_record_buffers(gss_engine *e, int n);
// Step graphs captured by the engine so far.
int64_t engine_graph_captures(const gss_engine *e);
int engine_serve_attach(gss_engine *e, const ServeReq *reqs, uint32_t req_cap,
                        const gss_var_patch *patches, uint64_t patch_cap, const uint64_t *head);
// gss_step_submit, or with serve a step that admits requests from the ring
// on the device (the descriptor then carries no loads, patches or valid
// slots of its own).
int step_submit_ex(gss_engine *e, const gss_step_desc *desc, bool serve);
// What the last waited serving step admitted.
int step_serve_report(gss_engine *e, ServeReport *out);
// For the last waited serving step: whether its records reached the host
// grouped by assignment word, and then offs[w] = first record of word w,
// offs[64] = all (65 entries).
int step_word_groups(gss_engine *e, uint64_t *offs, bool *grouped);
// Makes the engine's device current (GSS_OK or an error code).
int engine_activate(gss_engine *e);

}  // namespace gss
