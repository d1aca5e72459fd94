// Host-side helpers shared by the single-problem C ABI (hdg_api.hip) and the
// batches (hdg_batch.hip).  Not part of the public ABI.
#pragma once
#include <cstdint>
#include <vector>

#include "hdg_internal.h"

namespace hdg {

// Face end points fv[2f] < fv[2f+1] (0-based node ids), the vertex -> face
// lists in CSR form (faces in ascending order per vertex) and, when fq4 is
// given, the four vertices around each face (lower and higher end point,
// opposite vertex of cell 1 and of cell 2, that of cell 1 again on boundary
// faces).  faces and cells are the 1-based arrays of hdg_set_mesh.  Returns
// false when a face has no first cell.
bool face_vertex_maps(int64_t nn, int64_t nf, const int64_t* faces, const int64_t* cells, std::vector<int>& fv,
                      std::vector<int>& vf_ptr, std::vector<int>& vf_idx, std::vector<int>* fq4);

// Sizes of the vertex hierarchy of an nx x ny rectangle mesh: fills nlev,
// lev[1..nlev-1].nx, ny, n, npts and tail (no device arrays).
void mg_plan_levels(int nx, int ny, MgData& M);

}  // namespace hdg
