// Grid preparation on the device: Gaussian smoothing and resampling of velocity grids,
// the operations behind BaseGrid2D/3D.smooth and .resample (fteikpy/_base.py:81-123,
// :165-212), which the reference delegates to scipy.ndimage.gaussian_filter and
// scipy.interpolate.RegularGridInterpolator.  Both are restated from scipy's published
// algorithms (scipy 1.15), in the same operation order so that results match bit for bit:
//
// - gaussian_filter = one 1D correlation per axis (axis 0 first) with a normalised sampled
//   Gaussian of radius int(4 sigma + 0.5) and 'reflect' boundaries (d c b a | a b c d |
//   d c b a, repeated with period 2n); a symmetric kernel is applied as
//   in[i] w[0] + sum_{j=r..1} (in[i-j] + in[i+j]) w[j] (the weights come from the caller);
// - RegularGridInterpolator on the nodes of linspace(axis[0], axis[-1], m) per axis: cell
//   index by searchsorted(axis, q, 'right') - 1 clipped to [0, n-2], distance
//   y = (q - a[i]) / (a[i+1] - a[i]); 'linear' blends the 2^ndim corners (2D:
//   v00 (1-y0)(1-y1) + v01 (1-y0) y1 + v10 y0 (1-y1) + v11 y0 y1, left to right; 3D: corners
//   in lexicographic order, weight ((w0 w1) w2), summed in that order), 'nearest' takes
//   index i where y <= 0.5, else i + 1.
#include "fteik_amd.h"
#include "fteik_device.hpp"

#include <algorithm>
#include <vector>

namespace fteik {
namespace {

__device__ __forceinline__ long long reflect_index(long long i, long long n) {
    const long long p = 2 * n;
    i %= p;
    if (i < 0) i += p;
    return i < n ? i : p - 1 - i;
}

// One output element per thread; the line along `axis` has length n and stride `stride`;
// `inner` = product of the dimensions after `axis`.
__global__ void correlate1d_kernel(const double *in, double *out, long long total, int n,
                                   long long inner, const double *w, int radius) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const long long outer = e / (inner * n);
    const long long rem = e - outer * inner * n;
    const long long i = rem / inner, k = rem - i * inner;
    const double *line = in + outer * inner * n + k;
    double acc = line[i * inner] * w[radius];
    for (int j = radius; j >= 1; --j) {
        const double a = line[reflect_index(i - j, n) * inner];
        const double b = line[reflect_index(i + j, n) * inner];
        acc += (a + b) * w[radius - j];
    }
    out[e] = acc;
}

struct Resample {
    const double *in;
    double *out;
    const int *idx;   // per axis, concatenated: cell index
    const double *y;  // per axis, concatenated: distance in the cell
    int ndim, method;
    int n[3], m[3];   // input nodes, output nodes per axis
    int off[3];       // start of each axis in idx / y
};

__global__ void resample_kernel(const Resample R, long long total) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int q[3];
    long long r = e;
    for (int a = R.ndim - 1; a >= 0; --a) {
        q[a] = (int)(r % R.m[a]);
        r /= R.m[a];
    }
    int i[3];
    double y[3];
    for (int a = 0; a < R.ndim; ++a) {
        i[a] = R.idx[R.off[a] + q[a]];
        y[a] = R.y[R.off[a] + q[a]];
    }
    auto at = [&](int a0, int a1, int a2) {
        long long p = a0;
        p = p * R.n[1] + a1;
        if (R.ndim == 3) p = p * R.n[2] + a2;
        return R.in[p];
    };
    double v;
    if (R.method == 1) {  // nearest
        int j[3] = {0, 0, 0};
        for (int a = 0; a < R.ndim; ++a) j[a] = y[a] <= 0.5 ? i[a] : i[a] + 1;
        v = at(j[0], j[1], j[2]);
    } else if (R.ndim == 2) {
        const double y0 = y[0], y1 = y[1];
        v = at(i[0], i[1], 0) * (1.0 - y0) * (1.0 - y1) + at(i[0], i[1] + 1, 0) * (1.0 - y0) * y1 +
            at(i[0] + 1, i[1], 0) * y0 * (1.0 - y1) + at(i[0] + 1, i[1] + 1, 0) * y0 * y1;
    } else {
        v = 0.0;
        for (int c = 0; c < 8; ++c) {
            const int c0 = (c >> 2) & 1, c1 = (c >> 1) & 1, c2 = c & 1;
            double wt = c0 ? y[0] : 1.0 - y[0];
            wt = wt * (c1 ? y[1] : 1.0 - y[1]);
            wt = wt * (c2 ? y[2] : 1.0 - y[2]);
            v = v + at(i[0] + c0, i[1] + c1, i[2] + c2) * wt;
        }
    }
    R.out[e] = v;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace
}  // namespace fteik

using namespace fteik;

extern "C" int fteik_correlate1d(const double *in, double *out, int ndim, const int *shape,
                                 int axis, const double *weights, int radius, void *stream) {
    if (!in || !out || !shape || !weights || ndim < 1 || ndim > 3 || axis < 0 || axis >= ndim ||
        radius < 0 || in == out)
        return FTEIK_ERR_ARG;
    long long total = 1, inner = 1;
    for (int a = 0; a < ndim; ++a) {
        if (shape[a] < 1) return FTEIK_ERR_ARG;
        total *= shape[a];
        if (a > axis) inner *= shape[a];
    }
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(correlate1d_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, in,
                       out, total, shape[axis], inner, weights, radius);
    return hip_status(hipGetLastError());
}

extern "C" size_t fteik_resample_workspace_size(int ndim, const int *new_shape) {
    if (ndim < 2 || ndim > 3 || !new_shape) return 0;
    size_t m = 0;
    for (int a = 0; a < ndim; ++a) m += (size_t)std::max(new_shape[a], 0);
    return align256(sizeof(int) * m) + align256(sizeof(double) * m);
}

extern "C" int fteik_resample(const double *in, int ndim, const int *shape, const double *origin,
                              const double *gridsize, const int *new_shape, int method,
                              double *out, void *workspace, size_t workspace_bytes,
                              void *stream_) {
    if (!in || !out || !shape || !origin || !gridsize || !new_shape || !workspace ||
        ndim < 2 || ndim > 3 || (method != 0 && method != 1))
        return FTEIK_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (shape[a] < 2 || new_shape[a] < 1) return FTEIK_ERR_ARG;
    if (workspace_bytes < fteik_resample_workspace_size(ndim, new_shape)) return FTEIK_ERR_WORKSPACE;
    hipStream_t stream = (hipStream_t)stream_;
    Resample R;
    R.in = in;
    R.out = out;
    R.ndim = ndim;
    R.method = method;
    std::vector<int> idx;
    std::vector<double> ys;
    long long total = 1;
    for (int a = 0; a < 3; ++a) {
        R.n[a] = a < ndim ? shape[a] : 1;
        R.m[a] = a < ndim ? new_shape[a] : 1;
        R.off[a] = (int)idx.size();
        if (a >= ndim) continue;
        total *= new_shape[a];
        // axis: origin + d * arange(n) (fteikpy/_base.py:214-227); queries: numpy's
        // linspace(axis[0], axis[-1], m) = k * ((stop - start) / (m - 1)) + start, last = stop
        const int n = shape[a], m = new_shape[a];
        std::vector<double> ax(n);
        for (int k = 0; k < n; ++k) ax[k] = origin[a] + gridsize[a] * (double)k;
        const double start = ax[0], stop = ax[n - 1];
        const double step = m > 1 ? (stop - start) / (double)(m - 1) : 0.0;
        for (int k = 0; k < m; ++k) {
            double q = m > 1 ? (double)k * step + start : start;
            if (m > 1 && k == m - 1) q = stop;
            int i = (int)(std::upper_bound(ax.begin(), ax.end(), q) - ax.begin()) - 1;
            i = std::min(std::max(i, 0), n - 2);
            idx.push_back(i);
            ys.push_back((q - ax[i]) / (ax[i + 1] - ax[i]));
        }
    }
    int *idx_dev = (int *)workspace;
    double *y_dev = (double *)((char *)workspace + align256(sizeof(int) * idx.size()));
    R.idx = idx_dev;
    R.y = y_dev;
    FTEIK_CHECK(hipMemcpyAsync(idx_dev, idx.data(), sizeof(int) * idx.size(),
                               hipMemcpyHostToDevice, stream));
    FTEIK_CHECK(hipMemcpyAsync(y_dev, ys.data(), sizeof(double) * ys.size(),
                               hipMemcpyHostToDevice, stream));
    // the host vectors go out of scope on return: finish the copies first
    FTEIK_CHECK(hipStreamSynchronize(stream));
    hipLaunchKernelGGL(resample_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, R, total);
    return hip_status(hipGetLastError());
}
