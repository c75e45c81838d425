// Sweep history of a logged 3D solve: every node change (node, time before the change) in
// the order of the passes, and where each pass's changes begin.  Written by
// fteik3d_solve_logged, read (and rolled back) by fteik3d_adjoint_history.
//
// Layout of the caller's buffer (all offsets 256-byte aligned):
//   HistHeader | seg[max_pass + 1] (u64: entries before pass P) | node[cap] (u32: source *
//   nodes + node, or 0xffffffff for an unused entry) | old[cap] (f64)
// Writers reserve entries in chunks, so a pass's segment holds unused entries too.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace fteik {

struct HistHeader {
    unsigned long long count;  // changes logged (may exceed cap: then the log is incomplete)
    int npass;                 // passes recorded
    int max_pass;
    long long cap;             // entries the buffer holds
    long long nodes;           // nodes per source grid of the solve that wrote it
    int nsrc;
    int pad;
};

struct HistView {
    HistHeader *hdr;
    unsigned long long *seg;
    unsigned *node;
    double *old;
};

inline size_t hist_align(size_t n) { return (n + 255) & ~(size_t)255; }

inline size_t hist_bytes(long long cap, int max_pass) {
    return hist_align(sizeof(HistHeader)) + hist_align(sizeof(unsigned long long) * (max_pass + 1)) +
           hist_align(sizeof(unsigned) * (size_t)cap) + hist_align(sizeof(double) * (size_t)cap);
}

// Largest capacity a buffer of `bytes` holds for max_pass passes (0 if none): the inverse
// of hist_bytes.
inline long long hist_capacity(size_t bytes, int max_pass) {
    const size_t fixed = hist_align(sizeof(HistHeader)) +
                         hist_align(sizeof(unsigned long long) * (max_pass + 1));
    if (bytes <= fixed) return 0;
    long long cap = (long long)((bytes - fixed) / (sizeof(unsigned) + sizeof(double)));
    while (cap > 0 && hist_bytes(cap, max_pass) > bytes) --cap;  // alignment padding
    return cap;
}

inline HistView hist_view(void *base, long long cap, int max_pass) {
    char *p = (char *)base;
    HistView v;
    v.hdr = (HistHeader *)p;
    p += hist_align(sizeof(HistHeader));
    v.seg = (unsigned long long *)p;
    p += hist_align(sizeof(unsigned long long) * (max_pass + 1));
    v.node = (unsigned *)p;
    p += hist_align(sizeof(unsigned) * (size_t)cap);
    v.old = (double *)p;
    return v;
}

}  // namespace fteik
