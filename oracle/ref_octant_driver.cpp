// Test-only driver around the reference's own octant algebra
// (src/TMROctant.cpp), compiled from the reference sources by
// oracle/build.py into oracle/_ref/.  It exposes TMROctantArray::sort and
// TMROctant::childId so the oracle's restated Morton order can be pinned
// against the real implementation.
#include <cstring>

#include "TMROctant.h"

extern "C" {

// octs: n records of {block, x, y, z, tag, level, info} (24 bytes each,
// TMROctant layout).  Sorts (and de-duplicates) in place like
// TMROctantArray::sort; returns the new size.
int ref_sort_octants(void *octs, int n, int use_node_index) {
  TMROctant *arr = new TMROctant[n > 0 ? n : 1];
  std::memcpy(arr, octs, sizeof(TMROctant) * n);
  TMROctantArray list(arr, n, use_node_index);
  list.sort();
  TMROctant *out;
  int size;
  list.getArray(&out, &size);
  std::memcpy(octs, out, sizeof(TMROctant) * size);
  return size;
}

int ref_child_id(const void *oct) {
  TMROctant o;
  std::memcpy(&o, oct, sizeof(TMROctant));
  return o.childId();
}

int ref_compare(const void *a, const void *b) {
  TMROctant x, y;
  std::memcpy(&x, a, sizeof(TMROctant));
  std::memcpy(&y, b, sizeof(TMROctant));
  return x.compare(&y);
}

int ref_sizeof_octant(void) { return (int)sizeof(TMROctant); }
}
