// Server lifecycle on one MI355X: preprocess -> setup_server ->
// answer / answer_batch -> server_caps -> free_server.
//
// Drop-in for the reference's src/gpu_protocol.h (GpuServerConfig :22-24,
// GpuServerCaps :28-33, gpu_preprocess :44, gpu_setup_server :51-54,
// gpu_server_caps :56, gpu_answer :61, gpu_answer_batch :70-75,
// gpu_free_server :78). Unlike the reference header this one exposes no
// device-runtime types; errors are reported by exceptions
// (std::invalid_argument for bad input, std::runtime_error for device
// failures) and nothing is written when validation fails.
#pragma once
#include "protocol.h"

namespace inspire {

struct GpuServerCtx;  // opaque

struct GpuServerConfig {
    size_t max_batch = 16;  // answer_batch accepts at most this many queries
};

struct GpuServerCaps {
    size_t max_batch;          // slots allocated at setup
    size_t resident_bytes;     // DB + precomp + slot pools + tables
    size_t device_free_bytes;  // current free device memory
    size_t num_cts;            // ciphertexts per response
};

// db: db_rows*db_cols values in [0, P), row-major (db[row*db_cols + col]).
PreprocessData gpu_preprocess(const PublicParams& pp, const uint16_t* db);

GpuServerCtx* gpu_setup_server(const PublicParams& pp, const PreprocessData& precomp,
                               const GpuServerConfig& cfg);
GpuServerCtx* gpu_setup_server(const PublicParams& pp, const PreprocessData& precomp);

GpuServerCaps gpu_server_caps(const GpuServerCtx* ctx);

// Responses are in NTT form, one RlweCt per response ciphertext.
std::vector<RlweCt> gpu_answer(GpuServerCtx* ctx, const QueryMessage& qry);

// count <= max_batch independent queries; all-or-nothing (throws before any
// device work on invalid input). Results are bit-identical to count calls
// of gpu_answer.
std::vector<std::vector<RlweCt>> gpu_answer_batch(GpuServerCtx* ctx, const QueryMessage* queries,
                                                  size_t count);
inline std::vector<std::vector<RlweCt>> gpu_answer_batch(GpuServerCtx* ctx,
                                                         const std::vector<QueryMessage>& queries) {
    return gpu_answer_batch(ctx, queries.data(), queries.size());
}

void gpu_free_server(GpuServerCtx* ctx);

}  // namespace inspire
