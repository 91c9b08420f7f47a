// Admission of queued drag requests at an iteration boundary of the server (host only, no
// HIP): which requests join the in-flight set, and which meshes move between device and host
// memory under the HBM budget. Kept free of device state so the host tests run it directly.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace md {

struct AdmitMesh {
    size_t bytes = 0;        // device footprint while resident
    uint64_t last_use = 0;   // admission stamp of its latest request (least recently used first)
    bool resident = true;    // its arrays are on the device
    bool busy = false;       // a request on it is in flight
};

// Walks the queue (the mesh of each queued request, in submission order) and admits requests
// whose mesh is idle, at most max_active in flight (<= 0: no limit) counting `active`.
// A request on a non-resident mesh is admitted when its footprint fits `budget` (0: no limit)
// after evicting least recently used idle resident meshes; it waits otherwise, and so do later
// requests on non-resident meshes (first come, first served). A mesh chosen for eviction in
// this pass is not resident for the rest of the pass: a later request on it waits for the next
// boundary, when it is restored like any other evicted mesh.
// Out: take[q] = 1 for admitted queue entries, evict = meshes to move to host memory first,
// restore = meshes to upload (in that order); meshes' busy, last_use and resident flags,
// `resident_bytes` and `clock` are updated as if both lists were applied.
void plan_admission(std::vector<AdmitMesh> &meshes, const std::vector<int> &queue, int active, int max_active,
                    size_t budget, size_t &resident_bytes, uint64_t &clock, std::vector<char> &take,
                    std::vector<int> &evict, std::vector<int> &restore);

}  // namespace md
