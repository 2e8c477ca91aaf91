// Batched FFT convolution for long grids (gfx950).
//
// Every convolution of the objective is y = A x with A[n][n'] = s c[n-n'] + c[n+n'+1] (indices mod
// L = 2 NP), c the periodic IDFT of a sampled weight.  Extending x to length L as e[n'] = s x[n'],
// e[L-1-n'] = x[n'] turns A x into the circular convolution c (*) e, so with unnormalised DFTs
// (W = exp(-2 pi i / L), X the DFT of x zero-padded to L)
//     E[k] = s X[k] + W^-k conj(X[k]),     y = IDFT( sum_terms C_t[k] E_t[k] ) / L   on n < NP,
// which is how the reference evaluates it (Re IFFT(FFT(mirror(f)) w_hat), calc_rho_hat.jl:17-97),
// here with C the DFT of the same c the operator-GEMM applies.  For L = 2^q the transforms run as
// four-step FFTs, L = L1 x L2 with L2 = 256:
//   k_fft_cols_fwd  per input field: L1 length-L2 FFTs over n2 (x[n1 + L1 n2]), times W^(n1 k2),
//                   stored as rows k2 = 0..L2/2 of L1 values (real input: the other rows are the
//                   conjugates of these up to a phase and are never needed)
//   k_fft_mid       per point and block of rows: length-L1 FFTs of every input's rows give
//                   X[k2 + L2 k1]; each output column forms sum_t C_t E_t there, transforms back
//                   over k1 and applies W^-(n1 k2)
//   k_fft_cols_inv  per output column: the missing rows as conjugates (a real result has a
//                   Hermitian spectrum), length-L2 inverse FFTs over k2, y[n1 + L1 n2] for n < NP,
//                   stored with the job's epilogue (conv_store)
// Sub-transforms are radix-2 decimation in time in LDS (bit-reversed load, natural-order result).
// The path is an autotuner candidate beside the operator-GEMM tile shapes (CONV_FFT).
#include <cstdint>

#include "sdft_launch.h"

namespace sdft {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

__device__ __forceinline__ int brev_bits(int v, int lg) { return (int)(__brev((unsigned)v) >> (32 - lg)); }

// twr/twi[m] = W_N^m = W_L^(m L / N), m < N / 2
template <typename T>
__device__ __forceinline__ void load_twiddles(T* twr, T* twi, const cpx<T>* __restrict__ W, int N, int L) {
    const int step = L / N;
    for (int m = threadIdx.x; m < N / 2; m += 256) {
        const cpx<T> w = W[(int64_t)m * step];
        twr[m] = w.x;
        twi[m] = w.y;
    }
}

// 16 interleaved transforms of length N = 2^lgN (element (pos, col) at pos * 16 + col), input in
// bit-reversed order, output in natural order; INV uses the conjugate twiddles (no 1/N)
template <typename T, bool INV>
__device__ __forceinline__ void fft_cols16(T* re, T* im, const T* twr, const T* twi, int lgN) {
    const int N = 1 << lgN;
    for (int s = 1; s <= lgN; ++s) {
        const int half = 1 << (s - 1), step = N >> s;
        for (int j = threadIdx.x; j < 8 * N; j += 256) {
            const int col = j & 15, b = j >> 4;
            const int pos = b & (half - 1), grp = b >> (s - 1);
            const int i0 = (((grp << s) + pos) << 4) + col, i1 = i0 + (half << 4);
            const T wr = twr[pos * step], wi = INV ? -twi[pos * step] : twi[pos * step];
            const T xr = re[i1], xi = im[i1];
            const T br = xr * wr - xi * wi, bi = xr * wi + xi * wr;
            const T ar = re[i0], ai = im[i0];
            re[i0] = ar + br;
            im[i0] = ai + bi;
            re[i1] = ar - br;
            im[i1] = ai - bi;
        }
        __syncthreads();
    }
}

// nrows contiguous transforms of length N = 2^lgN (element (row, pos) at row * N + pos)
template <typename T, bool INV>
__device__ __forceinline__ void fft_rows(T* re, T* im, const T* twr, const T* twi, int lgN, int nrows) {
    const int N = 1 << lgN, hN = N >> 1;
    for (int s = 1; s <= lgN; ++s) {
        const int half = 1 << (s - 1), step = N >> s;
        for (int j = threadIdx.x; j < nrows * hN; j += 256) {
            const int row = j >> (lgN - 1), b = j & (hN - 1);
            const int pos = b & (half - 1), grp = b >> (s - 1);
            const int i0 = row * N + (grp << s) + pos, i1 = i0 + half;
            const T wr = twr[pos * step], wi = INV ? -twi[pos * step] : twi[pos * step];
            const T xr = re[i1], xi = im[i1];
            const T br = xr * wr - xi * wi, bi = xr * wi + xi * wr;
            const T ar = re[i0], ai = im[i0];
            re[i0] = ar + br;
            im[i0] = ai + bi;
            re[i1] = ar - br;
            im[i1] = ai - bi;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_fft_twiddles(cpx<double>* __restrict__ W64, cpx<float>* __restrict__ W32,
                                                      int L) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= L) return;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)L, &sn, &cs);  // 2 m / L exact: L is a power of two
    cpx<double> w;
    w.x = cs;
    w.y = -sn;
    W64[m] = w;
    cpx<float> w32;
    w32.x = (float)cs;
    w32.y = (float)-sn;
    W32[m] = w32;
}

template <typename T>
struct FftColsFwdArgs {
    const int32_t* pts;        // stage mode: point of blockIdx.z
    const FftSlot* slots;      // stage mode: input field of blockIdx.y
    int nslot;
    const T* rhoT;
    int Frt;
    const T* dphi;
    int Fdp;
    const T* raw;              // kernel-spectrum mode: vector blockIdx.y at raw + blockIdx.y * L
    const cpx<T>* W;
    cpx<T>* out;               // [vector][H][L1]
    int L, L1, L2, lg2, H;
};

// grid (L1 / 16, inputs, points); 16 consecutive n1 per workgroup
template <typename T>
__global__ void __launch_bounds__(256) k_fft_cols_fwd(FftColsFwdArgs<T> a, const DevPoint* __restrict__ pts,
                                                      const int32_t* __restrict__ evalflag, int prec) {
    const T* x;
    int nvalid;
    int64_t vec;
    if (a.raw != nullptr) {
        x = a.raw + (int64_t)blockIdx.y * a.L;
        nvalid = a.L;
        vec = blockIdx.y;
    } else {
        const int p = a.pts[blockIdx.z];
        if (evalflag[p] != prec) return;
        const DevPoint& P = pts[p];
        const FftSlot sl = a.slots[blockIdx.y];
        x = sl.infam == 0 ? a.rhoT + P.base * (int64_t)a.Frt + (int64_t)sl.infield * P.NP
                          : a.dphi + P.base * (int64_t)a.Fdp + (int64_t)sl.infield * P.NP;
        nvalid = P.NP;
        vec = (int64_t)blockIdx.z * a.nslot + blockIdx.y;
    }
    extern __shared__ unsigned char fft_smem[];
    T* re = (T*)fft_smem;
    T* im = re + 16 * a.L2;
    T* twr = im + 16 * a.L2;
    T* twi = twr + a.L2 / 2;
    const int n1_0 = blockIdx.x * 16;
    load_twiddles(twr, twi, a.W, a.L2, a.L);
    for (int e = threadIdx.x; e < 16 * a.L2; e += 256) {
        const int col = e & 15, n2 = e >> 4;
        const int n = n1_0 + col + a.L1 * n2;
        const int q = (brev_bits(n2, a.lg2) << 4) + col;
        re[q] = n < nvalid ? x[n] : T(0);
        im[q] = T(0);
    }
    __syncthreads();
    fft_cols16<T, false>(re, im, twr, twi, a.lg2);
    cpx<T>* o = a.out + vec * a.H * (int64_t)a.L1;
    for (int e = threadIdx.x; e < 16 * a.H; e += 256) {
        const int col = e & 15, k2 = e >> 4, n1 = n1_0 + col;
        const cpx<T> w = a.W[n1 * k2];  // n1 k2 < L1 L2 / 2
        const T vr = re[(k2 << 4) + col], vi = im[(k2 << 4) + col];
        cpx<T> r;
        r.x = vr * w.x - vi * w.y;
        r.y = vr * w.y + vi * w.x;
        o[(int64_t)k2 * a.L1 + n1] = r;
    }
}

// kernel spectra: rows of the column pass transformed over n1, natural order [vector][H][L1]
// (grid: (ceil(H / R), vectors))
__global__ void __launch_bounds__(256) k_fft_rows_fwd(const cpx<double>* __restrict__ sa, cpx<double>* __restrict__ chat,
                                                      const cpx<double>* __restrict__ W, int L, int L1, int lg1, int H,
                                                      int R) {
    extern __shared__ unsigned char fft_smem[];
    const int RL = R * L1;
    double* xr = (double*)fft_smem;
    double* xi = xr + RL;
    double* twr = xi + RL;
    double* twi = twr + L1 / 2;
    const int k2_0 = blockIdx.x * R;
    const int64_t vb = (int64_t)blockIdx.y * H * L1;
    load_twiddles(twr, twi, W, L1, L);
    for (int e = threadIdx.x; e < RL; e += 256) {
        const int r = e >> lg1, n1 = e & (L1 - 1), k2 = k2_0 + r;
        cpx<double> v;
        v.x = 0.0;
        v.y = 0.0;
        if (k2 < H) v = sa[vb + (int64_t)k2 * L1 + n1];
        const int q = r * L1 + brev_bits(n1, lg1);
        xr[q] = v.x;
        xi[q] = v.y;
    }
    __syncthreads();
    fft_rows<double, false>(xr, xi, twr, twi, lg1, R);
    for (int e = threadIdx.x; e < RL; e += 256) {
        const int r = e >> lg1, k1 = e & (L1 - 1), k2 = k2_0 + r;
        if (k2 >= H) continue;
        cpx<double> v;
        v.x = xr[e];
        v.y = xi[e];
        chat[vb + (int64_t)k2 * L1 + k1] = v;
    }
}

template <typename T>
struct FftMidArgs {
    const int32_t* pts;
    int nslot, njob;
    const FftJob* jobs;
    const cpx<T>* chat;   // [K][H][L1]
    const cpx<T>* W;
    const cpx<T>* sa;     // [point][slot][H][L1]
    cpx<T>* sb;           // [point][job][H][L1]
    int L, L1, L2, lg1, H, R;
};

// grid (ceil(H / R), points); R rows of L1 per workgroup (R L1 = 256)
template <typename T>
__global__ void __launch_bounds__(256) k_fft_mid(FftMidArgs<T> a, const int32_t* __restrict__ evalflag, int prec) {
    const int p = a.pts[blockIdx.y];
    if (evalflag[p] != prec) return;
    extern __shared__ unsigned char fft_smem[];
    const int RL = a.R * a.L1;
    T* xr = (T*)fft_smem;
    T* xi = xr + a.nslot * RL;
    T* yr = xi + a.nslot * RL;
    T* yi = yr + RL;
    T* twr = yi + RL;
    T* twi = twr + a.L1 / 2;
    const int k2_0 = blockIdx.x * a.R;
    load_twiddles(twr, twi, a.W, a.L1, a.L);
    const cpx<T>* in = a.sa + (int64_t)blockIdx.y * a.nslot * a.H * a.L1;
    for (int e = threadIdx.x; e < a.nslot * RL; e += 256) {
        const int s = e / RL, rem = e - s * RL;
        const int r = rem >> a.lg1, n1 = rem & (a.L1 - 1), k2 = k2_0 + r;
        cpx<T> v;
        v.x = T(0);
        v.y = T(0);
        if (k2 < a.H) v = in[((int64_t)s * a.H + k2) * a.L1 + n1];
        const int q = s * RL + r * a.L1 + brev_bits(n1, a.lg1);
        xr[q] = v.x;
        xi[q] = v.y;
    }
    __syncthreads();
    fft_rows<T, false>(xr, xi, twr, twi, a.lg1, a.nslot * a.R);
    cpx<T>* out = a.sb + (int64_t)blockIdx.y * a.njob * a.H * a.L1;
    for (int j = 0; j < a.njob; ++j) {
        const FftJob& J = a.jobs[j];  // read in place: term fields at constant offsets (no private copy)
        const int nterms = J.nterms;
        for (int e = threadIdx.x; e < RL; e += 256) {
            const int r = e >> a.lg1, k1 = e & (a.L1 - 1), k2 = k2_0 + r;
            T accr = T(0), acci = T(0);
            if (k2 < a.H) {
                const cpx<T> w = a.W[k2 + a.L2 * k1];
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    if (t >= nterms) break;
                    const int q = J.slot[t] * RL + r * a.L1 + k1;
                    const T Xr = xr[q], Xi = xi[q];
                    const T sg = J.sign[t] < 0 ? T(-1) : T(1);
                    // E = s X + conj(W^k X)
                    const T wxr = w.x * Xr - w.y * Xi, wxi = w.x * Xi + w.y * Xr;
                    const T Er = sg * Xr + wxr, Ei = sg * Xi - wxi;
                    const cpx<T> ch = a.chat[((int64_t)J.kern[t] * a.H + k2) * a.L1 + k1];
                    accr += ch.x * Er - ch.y * Ei;
                    acci += ch.x * Ei + ch.y * Er;
                }
            }
            const int q = r * a.L1 + brev_bits(k1, a.lg1);
            yr[q] = accr;
            yi[q] = acci;
        }
        __syncthreads();
        fft_rows<T, true>(yr, yi, twr, twi, a.lg1, a.R);
        for (int e = threadIdx.x; e < RL; e += 256) {
            const int r = e >> a.lg1, n1 = e & (a.L1 - 1), k2 = k2_0 + r;
            if (k2 >= a.H) continue;
            const cpx<T> w = a.W[n1 * k2];
            const T vr = yr[e], vi = yi[e];
            cpx<T> o;
            o.x = vr * w.x + vi * w.y;  // times conj(W^(n1 k2))
            o.y = vi * w.x - vr * w.y;
            out[((int64_t)j * a.H + k2) * a.L1 + n1] = o;
        }
        __syncthreads();
    }
}

template <typename T>
struct FftColsInvArgs {
    const int32_t* pts;
    const FftJob* fjobs;
    int njob;
    const ConvJob* jobs;
    const cpx<T>* W;
    const cpx<T>* sb;
    T* nw;
    int Fnw;
    double* fx;
    double* wsum;
    int Fws;
    int L, L1, L2, lg2, H;
    double invL;
};

// grid (L1 / 16, jobs, points)
template <typename T>
__global__ void __launch_bounds__(256) k_fft_cols_inv(FftColsInvArgs<T> a, const DevPoint* __restrict__ pts,
                                                      const int32_t* __restrict__ evalflag, int prec) {
    const int p = a.pts[blockIdx.z];
    if (evalflag[p] != prec) return;
    extern __shared__ unsigned char fft_smem[];
    T* re = (T*)fft_smem;
    T* im = re + 16 * a.L2;
    T* twr = im + 16 * a.L2;
    T* twi = twr + a.L2 / 2;
    const int n1_0 = blockIdx.x * 16;
    load_twiddles(twr, twi, a.W, a.L2, a.L);
    const cpx<T>* v = a.sb + ((int64_t)blockIdx.z * a.njob + blockIdx.y) * a.H * (int64_t)a.L1;
    for (int e = threadIdx.x; e < 16 * a.L2; e += 256) {
        const int col = e & 15, k2 = e >> 4, n1 = n1_0 + col;
        cpx<T> z;
        if (k2 < a.H) {
            z = v[(int64_t)k2 * a.L1 + n1];
        } else {  // row L2 - k2 conjugated (Hermitian spectrum of a real result)
            z = v[(int64_t)(a.L2 - k2) * a.L1 + n1];
            z.y = -z.y;
        }
        const int q = (brev_bits(k2, a.lg2) << 4) + col;
        re[q] = z.x;
        im[q] = z.y;
    }
    __syncthreads();
    fft_cols16<T, true>(re, im, twr, twi, a.lg2);
    const DevPoint& P = pts[p];
    const ConvJob& J = a.jobs[a.fjobs[blockIdx.y].job];
    for (int e = threadIdx.x; e < 8 * a.L2; e += 256) {  // n2 < L2 / 2: n < L / 2 = NP
        const int col = e & 15, n2 = e >> 4;
        const int n = n1_0 + col + a.L1 * n2;
        conv_store<T>(J, P, n, (double)re[(n2 << 4) + col] * a.invL, a.nw, a.Fnw, a.fx, a.wsum, a.Fws);
    }
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
static int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

bool fft_length_ok(int L) { return L >= 4096 && L <= 65536 && (L & (L - 1)) == 0; }

void fft_split(int L, int* L1, int* L2) {
    *L2 = 256;
    *L1 = L / 256;
}

int64_t fft_spectrum_len(int L) {
    int L1, L2;
    fft_split(L, &L1, &L2);
    return (int64_t)(L2 / 2 + 1) * L1;
}

static size_t cols_lds_bytes(int L2, size_t elem) { return (size_t)(32 * L2 + L2) * elem; }

size_t fft_mid_lds_bytes(int nslot, int L1, size_t elem) { return (size_t)((2 * nslot + 2) * 256 + L1) * elem; }

hipError_t fft_prepare_attributes() {
    const int mx = 160 * 1024;
    const void* fns[] = {reinterpret_cast<const void*>(&k_fft_cols_fwd<float>),
                         reinterpret_cast<const void*>(&k_fft_cols_fwd<double>),
                         reinterpret_cast<const void*>(&k_fft_mid<float>),
                         reinterpret_cast<const void*>(&k_fft_mid<double>),
                         reinterpret_cast<const void*>(&k_fft_cols_inv<float>),
                         reinterpret_cast<const void*>(&k_fft_cols_inv<double>),
                         reinterpret_cast<const void*>(&k_fft_rows_fwd)};
    for (const void* f : fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, mx);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

void launch_fft_twiddles(hipStream_t st, int L, void* W64, void* W32) {
    hipLaunchKernelGGL(k_fft_twiddles, dim3((L + 255) / 256), dim3(256), 0, st, (cpx<double>*)W64, (cpx<float>*)W32, L);
}

// C_k = DFT(c_k) of the K kernel vectors of one set (c64: K vectors of length L), [K][H][L1] in
// FP64 and FP32; tmp holds K spectra of the column pass
void launch_fft_kernel_spectra(hipStream_t st, int L, int K, const double* c64, const void* W64, void* tmp,
                               void* chat64, void* chat32) {
    int L1, L2;
    fft_split(L, &L1, &L2);
    const int H = L2 / 2 + 1, R = 256 / L1;
    FftColsFwdArgs<double> a;
    a.pts = nullptr;
    a.slots = nullptr;
    a.nslot = 0;
    a.rhoT = nullptr;
    a.Frt = 0;
    a.dphi = nullptr;
    a.Fdp = 0;
    a.raw = c64;
    a.W = (const cpx<double>*)W64;
    a.out = (cpx<double>*)tmp;
    a.L = L;
    a.L1 = L1;
    a.L2 = L2;
    a.lg2 = ilog2(L2);
    a.H = H;
    hipLaunchKernelGGL(k_fft_cols_fwd<double>, dim3(L1 / 16, K, 1), dim3(256), cols_lds_bytes(L2, 8), st, a,
                       (const DevPoint*)nullptr, (const int32_t*)nullptr, 0);
    hipLaunchKernelGGL(k_fft_rows_fwd, dim3((H + R - 1) / R, K), dim3(256), (size_t)(2 * 256 + L1) * 8, st,
                       (const cpx<double>*)tmp, (cpx<double>*)chat64, (const cpx<double>*)W64, L, L1, ilog2(L1), H, R);
    launch_convert_f32(st, (const double*)chat64, (float*)chat32, 2 * (int64_t)K * H * L1);
}

template <typename T>
void launch_fft_conv(hipStream_t st, const Ctx& c, int prec, const OpPlan& plan, const T* rhoT, const T* dphi,
                     T* nw, double* fx) {
    const int w = sizeof(T) == 8 ? 1 : 0;
    for (const FftSetPlan& f : plan.fft) {
        const FftRes& r = c.fft_res[f.kset];
        const cpx<T>* W = (const cpx<T>*)(w ? r.W64 : r.W32);
        const cpx<T>* chat = (const cpx<T>*)(w ? r.chat64 : r.chat32);
        cpx<T>* sa = (cpx<T>*)c.fft_sa[w];
        cpx<T>* sb = (cpx<T>*)c.fft_sb[w];
        FftColsFwdArgs<T> a;
        a.pts = f.d_pts;
        a.slots = f.d_slots;
        a.nslot = f.nslot;
        a.rhoT = rhoT;
        a.Frt = c.Frt;
        a.dphi = dphi;
        a.Fdp = c.Fdp;
        a.raw = nullptr;
        a.W = W;
        a.out = sa;
        a.L = f.L;
        a.L1 = f.L1;
        a.L2 = f.L2;
        a.lg2 = f.lg2;
        a.H = f.H;
        hipLaunchKernelGGL(k_fft_cols_fwd<T>, dim3(f.L1 / 16, f.nslot, f.npts), dim3(256), cols_lds_bytes(f.L2, sizeof(T)),
                           st, a, c.d_pts, c.d_evalflag, prec);
        FftMidArgs<T> m;
        m.pts = f.d_pts;
        m.nslot = f.nslot;
        m.njob = f.njob;
        m.jobs = f.d_jobs;
        m.chat = chat;
        m.W = W;
        m.sa = sa;
        m.sb = sb;
        m.L = f.L;
        m.L1 = f.L1;
        m.L2 = f.L2;
        m.lg1 = f.lg1;
        m.H = f.H;
        m.R = f.R;
        hipLaunchKernelGGL(k_fft_mid<T>, dim3((f.H + f.R - 1) / f.R, f.npts), dim3(256),
                           fft_mid_lds_bytes(f.nslot, f.L1, sizeof(T)), st, m, c.d_evalflag, prec);
        FftColsInvArgs<T> v;
        v.pts = f.d_pts;
        v.fjobs = f.d_jobs;
        v.njob = f.njob;
        v.jobs = c.d_jobs;
        v.W = W;
        v.sb = sb;
        v.nw = nw;
        v.Fnw = c.Fnw;
        v.fx = fx;
        v.wsum = c.d_wsum;
        v.Fws = c.Fws;
        v.L = f.L;
        v.L1 = f.L1;
        v.L2 = f.L2;
        v.lg2 = f.lg2;
        v.H = f.H;
        v.invL = 1.0 / (double)f.L;
        hipLaunchKernelGGL(k_fft_cols_inv<T>, dim3(f.L1 / 16, f.njob, f.npts), dim3(256),
                           cols_lds_bytes(f.L2, sizeof(T)), st, v, c.d_pts, c.d_evalflag, prec);
    }
}

template void launch_fft_conv<float>(hipStream_t, const Ctx&, int, const OpPlan&, const float*, const float*, float*,
                                     double*);
template void launch_fft_conv<double>(hipStream_t, const Ctx&, int, const OpPlan&, const double*, const double*,
                                      double*, double*);

}  // namespace sdft
