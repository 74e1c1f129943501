// Device side of the triangle-asterism and star-pair similarity registration
// (cpp/cuda/registration_kernels.cu:746-1821 semantics), shared by
// triangles.hip (kernels) and capi_triangles.cpp (C ABI, host ordering).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gh {

// One catalogue of a descriptor launch: raw slots of anchor a, combination c
// land at raw_offset + a * combos + c.
struct TriCatalog {
  const float* x;
  const float* y;
  int count;
  int neighbors;
  int combos;
  int raw_offset;
};

// A raw descriptor slot.  valid == 0 marks a slot some gate rejected.
struct RawTriangle {
  float d0, d1, area, pad;
  int first, second, third, valid;
};

// Best candidate of one scoring block (score -1: none).
struct BlockBest {
  int score;
  float rms;
  int index;
  int pos;
};

// Device result of one similarity estimate; the host copies it out as is.
struct SimilarityOut {
  float matrix[9];
  float scale, rotation, rms;
  int inliers, best_index, best_pos;
  int refit_inliers, refit_status;
  float refit_rms;
  unsigned scored;
  int pad;
};

// The two catalogues of one estimate.  *_all: the caller's arrays in their
// own order (descriptor / pair indices and the refit tree order refer to
// them); *_fin: their finite entries, order kept (the scoring loops).
struct ScoreCatalogs {
  const float2* ref_fin;
  int ref_fin_count;
  const float2* mov_fin;
  int mov_fin_count;
  const float* ref_x;
  const float* ref_y;
  int ref_count;
  const float* mov_x;
  const float* mov_y;
  int mov_count;
};

// Passing candidates: index, (a, b, tx, ty), then the scorer's inliers / RMS.
struct CandidateList {
  int* index;
  float4* params;
  int* score;
  float* rms;
  unsigned* count;
  unsigned capacity;
};

struct CatalogGates {
  float min_pair_distance, prior_dx, prior_dy, prior_radius_px, min_scale, max_scale, max_abs_rotation_rad;
};

// Zeroes the candidate counter (a kernel, so the sequence stays capturable).
void launch_reset_count(unsigned* count, hipStream_t stream);

void launch_triangle_descriptors(const TriCatalog* catalogs, int n_catalogs, int max_count, RawTriangle* raw,
                                 hipStream_t stream);

void launch_triangle_candidates(const ScoreCatalogs& cat, const float2* ref_desc, const int* ref_idx, int ref_n,
                                const float2* mov_desc, const int* mov_idx, int mov_n, float radius,
                                const CandidateList& out, hipStream_t stream);

void launch_pair_candidates(const ScoreCatalogs& cat, const CatalogGates& gates, long long candidate_count,
                            const CandidateList& out, hipStream_t stream);

// Scores every listed candidate (one wave each) and leaves per-block bests;
// returns the number of blocks written.
int launch_score_candidates(const ScoreCatalogs& cat, float tolerance, const CandidateList& list,
                            BlockBest* block_best, int max_blocks, hipStream_t stream);

// Best candidate, its matrix, and (refit != 0) the catalogue refit.
void launch_finalize(const ScoreCatalogs& cat, float tolerance, const CandidateList& list, const BlockBest* block_best,
                     int blocks, int refit, SimilarityOut* out, hipStream_t stream);

int score_max_blocks();

}  // namespace gh
