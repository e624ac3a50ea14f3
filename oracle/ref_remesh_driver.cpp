// Test driver (TEST INFRASTRUCTURE ONLY) around the reference's own soup
// splitters, compiled from the reference's sources where they lie
// (src/remesh.cpp, src/remesh_clip.cpp; see oracle/Makefile, target ref).
// It replaces the nanobind binding of src/main.cpp:308-385 with two C entry
// points that run a splitter and copy its four outputs into caller buffers,
// so the tests can compare the library's splitter with the reference's on
// the same inputs.  Nothing here is shipped or measured.
#include <vector>

#include "remesh.h"
#include "remesh_clip.h"

namespace {

template <class S>
int copy_out(const S& s, int cap_v, int cap_t, float* ov, int* of, int* omap, int* oflag, int* nv, int* nt) {
    *nv = s.getNumVertices();
    *nt = s.getNumTriangles();
    if (*nv > cap_v || *nt > cap_t) return 1;   // caller retries with the sizes returned
    s.exportToFlatArrays(ov, of);
    s.getFaceMapping(omap);
    s.getSameAsOriginal(oflag);
    return 0;
}

}  // namespace

extern "C" {

int ref_split_soup(const float* verts, int nv_in, const int* faces, int nt_in, int num_splits, float tau,
                   int cap_v, int cap_t, float* ov, int* of, int* omap, int* oflag, int* nv, int* nt) {
    diffsoup::TriangleSoupSplitter s(verts, faces, nv_in, nt_in);
    s.splitLongEdges(num_splits, tau);
    return copy_out(s, cap_v, cap_t, ov, of, omap, oflag, nv, nt);
}

int ref_split_soup_clip(const float* mvp, const float* verts, int nv_in, const int* faces, int nt_in,
                        const int* valid, int num_splits, float tau_ratio, float aspect_wh,
                        int cap_v, int cap_t, float* ov, int* of, int* omap, int* oflag, int* nv, int* nt) {
    diffsoup::TriangleSoupSplitterClip s(mvp, verts, faces, nv_in, nt_in, valid);
    s.splitLongEdges(num_splits, tau_ratio, aspect_wh);
    return copy_out(s, cap_v, cap_t, ov, of, omap, oflag, nv, nt);
}

}  // extern "C"
