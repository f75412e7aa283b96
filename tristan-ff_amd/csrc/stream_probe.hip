// stream_probe.hip -- the streaming bandwidth this GPU delivers for the
// stage kernel's mix of HBM streams, as a ceiling to read its roofline
// fraction against.
//
// A fused stage launch reads and writes whole field components (one array
// of n cells each): stage 1 reads the 6 raw components and writes 9 (the
// raw output and the cleaned U^n E), stages 2 and 3 read 12 (raw input and
// U^n) and write 6.  This program moves exactly those streams with no
// arithmetic, halos or LDS: every thread sums its element of each input
// array and stores the sum to each output array with a non-temporal store
// (as the stage kernel stores its outputs), grid-stride over n elements.
// It prints, per mix and word size, the time per launch and the bytes moved
// per second (the best of several rounds of timed launches).
//
//   stream_probe [n]      n cells per component (default 256^3)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int MAXS = 12;

template <class S>
struct Streams {
  const S* in[MAXS];
  S* out[MAXS];
};

template <class S, int NR, int NW>
__global__ void __launch_bounds__(256) k_streams(Streams<S> p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    S s = 0;
#pragma unroll
    for (int r = 0; r < NR; ++r) s += p.in[r][i];
#pragma unroll
    for (int w = 0; w < NW; ++w) __builtin_nontemporal_store(s + (S)w, &p.out[w][i]);
  }
}

bool check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "stream_probe: %s: %s\n", what, hipGetErrorString(e));
    return false;
  }
  return true;
}

// the best of `rounds` rounds of `reps` launches, in microseconds per launch
template <class S, int NR, int NW>
bool time_mix(const Streams<S>& p, long n, const char* name, const char* word) {
  const int grid = 16384, block = 256, reps = 10, rounds = 3;
  hipEvent_t a, b;
  if (!check(hipEventCreate(&a), "event") || !check(hipEventCreate(&b), "event")) return false;
  hipLaunchKernelGGL((k_streams<S, NR, NW>), dim3(grid), dim3(block), 0, 0, p, n);
  if (!check(hipDeviceSynchronize(), "warm-up launch")) return false;
  float best = 1e30f;
  for (int r = 0; r < rounds; ++r) {
    if (!check(hipEventRecord(a), "event record")) return false;
    for (int k = 0; k < reps; ++k) hipLaunchKernelGGL((k_streams<S, NR, NW>), dim3(grid), dim3(block), 0, 0, p, n);
    if (!check(hipEventRecord(b), "event record")) return false;
    if (!check(hipEventSynchronize(b), "timed launches")) return false;
    float ms = 0.f;
    if (!check(hipEventElapsedTime(&ms, a, b), "elapsed time")) return false;
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  const double us = 1e3 * best / reps;
  const double bytes = (double)(NR + NW) * n * sizeof(S);
  std::printf("| %s | %s | %d | %d | %.1f | %.1f | %.2f |\n", name, word, NR, NW, bytes / 1e6, us,
              bytes / (us * 1e-6) / 1e12);
  std::fflush(stdout);
  return true;
}

template <class S>
bool run_word(long n, const char* word) {
  // 12 input and 9 output arrays of n elements; each kernel reads and
  // writes elements 0..n-1 of the arrays it is given
  std::vector<S*> bufs;
  bool ok = true;
  for (int i = 0; i < 21 && ok; ++i) {
    S* q = nullptr;
    ok = check(hipMalloc(&q, (size_t)n * sizeof(S)), "hipMalloc") &&
         check(hipMemset(q, 0, (size_t)n * sizeof(S)), "hipMemset");
    if (q) bufs.push_back(q);
  }
  if (ok) {
    Streams<S> p{};
    for (int i = 0; i < MAXS; ++i) {
      p.in[i] = bufs[i];
      p.out[i] = bufs[12 + (i % 9)];
    }
    ok = time_mix<S, 6, 9>(p, n, "stage 1 (6 read, 9 written)", word) &&
         time_mix<S, 12, 6>(p, n, "stages 2 and 3 (12 read, 6 written)", word) &&
         time_mix<S, 3, 3>(p, n, "copy of one field (3 read, 3 written)", word);
  }
  for (S* q : bufs) (void)hipFree(q);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 256L * 256 * 256;
  if (n <= 0) {
    std::fprintf(stderr, "usage: stream_probe [cells per component > 0]\n");
    return 2;
  }
  std::printf("| streams | word | arrays read | arrays written | MB per launch | us per launch | TB/s |\n");
  std::printf("|---|---|---|---|---|---|---|\n");
  if (!run_word<float>(n, "FP32") || !run_word<double>(n, "FP64")) return 1;
  return 0;
}
