// clm_layout.h -- device layout of a column batch.
//
// Columns are sorted by layer count (descending) and dealt into packs of
// 64: one pack = one wavefront, one column per lane.  Inside a pack the
// cells are stored layer-major with the lane fastest,
//     cell(pack p, lane l, layer j) = cell_off[p] + 64 j + l,
//     face(pack p, lane l, face f)  = face_off[p] + 64 f + l,
// so that a wavefront walking its columns level by level reads and writes
// one contiguous 512-byte line per field and level (fully coalesced), and
// a pack is padded only up to its own deepest column.  "Slots" number the
// lanes of all packs (slot = 64 p + l); per-column data (parameters,
// boundary fluxes, ∫F) are stored per slot.  Padding lanes have nl = 0.
#pragma once
#include <stdint.h>

namespace clm {

constexpr int kWave = 64;

struct Layout {
    const int64_t* cell_off;  // [npack + 1]
    const int64_t* face_off;  // [npack + 1]
    const int32_t* nl;        // [nslot] layers of the slot's column (0 = padding)
    const int32_t* slot_col;  // [nslot] CSR column of the slot (-1 = padding)
    int32_t npack;
    int64_t nslot;
};

// per-lane view of one column
struct ColView {
    int64_t cb;  // cell base (layer j at cb + 64 j)
    int64_t fb;  // face base
    int32_t N;
    int64_t s;
};

}  // namespace clm
