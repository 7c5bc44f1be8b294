// Independent check of the owner-tile accumulation schedules (the race
// detection of the LDS scatter): written separately from the colouring in
// Problem::build_tiles and run on its output before upload.
//
// Vector-ALU records of a tile are processed in rounds, one colour phase at
// a time with a barrier between phases; they are race-free when
//   * the records are sorted by colour (the phases of a round are the
//     colours between its first and last record),
//   * no two records of one colour write the same LDS slot,
//   * every written corner is an accumulated slot (position < na).
// Matrix-core records are grouped per wave: the waves' written slots must be
// disjoint (no barriers in that element phase) and, within a wave, records
// of one colour must not share a slot (a 16-record batch writes its rows
// together).
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "kernels.h"

namespace tmrx {

int check_tile_schedule(int ntiles, const int *rec_ptr, const TileElem *recs, const int *nacc,
                        int waves, const int *wave_ptr, const TileElem *mrecs, std::string &err) {
  std::vector<int> seen;  // per slot position: 1 + the phase that last wrote it
  for (int t = 0; t < ntiles; t++) {
    const int na = nacc[t];
    const int npos = round8(na) + 8;
    seen.assign((size_t)npos * 1, 0);
    // vector-ALU records
    int prev_color = -1;
    std::vector<std::vector<int>> by_color;
    for (int k = rec_ptr[t]; k < rec_ptr[t + 1]; k++) {
      const TileElem &r = recs[k];
      if ((int)r.color < prev_color) {
        err = "tile " + std::to_string(t) + ": records not sorted by colour at " + std::to_string(k);
        return -1;
      }
      if ((int)r.color != prev_color) {
        prev_color = r.color;
        seen.assign((size_t)npos, 0);  // a new phase
      }
      for (int a = 0; a < 8; a++) {
        if (!(r.mask & (1u << a))) continue;
        const int s = tile_pos(r.lc[a]);
        if (s >= na) {
          err = "tile " + std::to_string(t) + ": record " + std::to_string(k) +
                " writes slot " + std::to_string(s) + " beyond the accumulators (" +
                std::to_string(na) + ")";
          return -1;
        }
        if (seen[s]) {
          err = "tile " + std::to_string(t) + ": colour " + std::to_string(r.color) +
                " writes slot " + std::to_string(s) + " twice (records " +
                std::to_string(seen[s] - 1) + " and " + std::to_string(k) + ")";
          return -1;
        }
        seen[s] = k + 1;
      }
    }
    if (!wave_ptr || !mrecs || waves <= 0) continue;
    // matrix-core records: per wave, slots owned by that wave only, and
    // one writer per slot within a (wave, colour) group
    std::vector<int> owner((size_t)npos, -1);
    for (int w = 0; w < waves; w++) {
      const int b0 = 16 * wave_ptr[(size_t)t * (waves + 1) + w];
      const int b1 = 16 * wave_ptr[(size_t)t * (waves + 1) + w + 1];
      int color = -1;
      std::vector<int> last((size_t)npos, -1);
      for (int k = b0; k < b1; k++) {
        const TileElem &r = mrecs[k];
        if ((int)r.color != color) {
          color = r.color;
          std::fill(last.begin(), last.end(), -1);
        }
        for (int a = 0; a < 8; a++) {
          if (!(r.mask & (1u << a))) continue;
          const int s = tile_pos(r.lc[a]);
          if (s >= na) {
            err = "tile " + std::to_string(t) + ": matrix-core record writes slot beyond the "
                  "accumulators";
            return -1;
          }
          if (owner[s] >= 0 && owner[s] != w) {
            err = "tile " + std::to_string(t) + ": slot " + std::to_string(s) +
                  " written by waves " + std::to_string(owner[s]) + " and " + std::to_string(w);
            return -1;
          }
          owner[s] = w;
          if (last[s] >= 0) {
            err = "tile " + std::to_string(t) + ": wave " + std::to_string(w) + " colour " +
                  std::to_string(color) + " writes slot " + std::to_string(s) + " twice";
            return -1;
          }
          last[s] = k;
        }
      }
    }
  }
  return 0;
}

}  // namespace tmrx

extern "C" int tmrx_check_tile_schedule(int ntiles, const int *rec_ptr, const void *recs,
                                        const int *nacc, char *msg, int msglen) {
  std::string err;
  const int r = tmrx::check_tile_schedule(ntiles, rec_ptr,
                                          static_cast<const tmrx::TileElem *>(recs), nacc, 0,
                                          nullptr, nullptr, err);
  if (msg && msglen > 0) {
    const size_t n = std::min((size_t)msglen - 1, err.size());
    err.copy(msg, n);
    msg[n] = 0;
  }
  return r;
}
