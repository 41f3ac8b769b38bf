// Host-side parts of liblapheig: error strings, the seeded synthetic gauge
// field, ILDG LIME ingest/egress, spectrum recovery, phase fixing, the
// tr(V^H V) guard and the eigensystem writers.  These are O(file) or
// O(n * nev) and run on the CPU beside the device solve, exactly where the
// reference runs them (ev_ts.cpp:170-171, 304-319).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <thread>
#include <vector>

#include "common.h"

namespace laph {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// Splits [0, n) over up to 16 host threads.
template <class F>
static void parallel_for(int64_t n, F&& f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> pool;
  int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back([&f, lo, hi] {
      for (int64_t i = lo; i < hi; ++i) f(i);
    });
  }
  for (auto& th : pool) th.join();
}

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Counter-based RNG: splitmix64 finaliser applied to seed*golden + (ctr+1)*odd.
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double hash_u01(uint64_t seed, uint64_t ctr) {
  uint64_t h = seed * 0x9E3779B97F4A7C15ull + (ctr + 1) * 0xD1B54A32D192ED03ull;
  return ((double)(mix64(h) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Hermitian 3x3 eigen-decomposition by cyclic complex Jacobi rotations.
static void herm3_eig(cd A[9], double w[3], cd Q[9]) {
  for (int i = 0; i < 9; ++i) Q[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = std::abs(A[1]) + std::abs(A[2]) + std::abs(A[5]);
    double diag = std::abs(A[0].real()) + std::abs(A[4].real()) + std::abs(A[8].real());
    if (off <= 1e-300 || off <= 1e-18 * diag) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double g = std::abs(A[3 * p + q]);
        if (g == 0.0) continue;
        cd ph = A[3 * p + q] / g;
        double tau = (A[4 * q].real() - A[4 * p].real()) / (2.0 * g);
        double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        cd Jqp = -s * std::conj(ph), Jqq = c * std::conj(ph);
        for (int k = 0; k < 3; ++k) {  // A <- A J (columns p, q)
          cd akp = A[3 * k + p], akq = A[3 * k + q];
          A[3 * k + p] = akp * c + akq * Jqp;
          A[3 * k + q] = akp * s + akq * Jqq;
          cd qkp = Q[3 * k + p], qkq = Q[3 * k + q];
          Q[3 * k + p] = qkp * c + qkq * Jqp;
          Q[3 * k + q] = qkp * s + qkq * Jqq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^H A (rows p, q)
          cd apk = A[3 * p + k], aqk = A[3 * q + k];
          A[3 * p + k] = c * apk + std::conj(Jqp) * aqk;
          A[3 * q + k] = s * apk + std::conj(Jqq) * aqk;
        }
        A[3 * p + q] = A[3 * q + p] = 0.0;
      }
  }
  for (int i = 0; i < 3; ++i) w[i] = A[4 * i].real();
}

}  // namespace laph

using laph::cd;

extern "C" {

const char* laph_error_string(int code) {
  switch (code) {
    case LAPH_OK: return "ok";
    case LAPH_ERR_INVALID: return "invalid argument";
    case LAPH_ERR_HIP: return "HIP runtime or library failure";
    case LAPH_ERR_NOMEM: return "out of memory";
    case LAPH_ERR_IO: return "file I/O failure";
    case LAPH_ERR_FORMAT: return "LIME record missing or wrong size";
    case LAPH_ERR_NOCONV: return "eigensolver did not converge";
    case LAPH_ERR_TRACE: return "tr(V^H V) guard failed";
    default: return "unknown error";
  }
}

const char* laph_last_error(void) { return laph::g_last_error.c_str(); }

const char* laph_version(void) { return "lapheig-mi355x 0.1 (gfx950)"; }

int laph_synth_config(unsigned long long seed, double eps, int L, int T, double* config) {
  LAPH_CHECK(L > 0 && T > 0 && config, LAPH_ERR_INVALID, "laph_synth_config: bad arguments");
  return laph::guarded([&] {
    const int64_t nlinks = (int64_t)T * L * L * L * 4;
    laph::parallel_for(nlinks, [&](int64_t c) {
      // c = ((((t*L+x)*L+y)*L+z)*4 + mu), mu in ILDG order x,y,z,t
      double xi[8];
      for (int j = 0; j < 8; ++j) {
        double u1 = laph::hash_u01(seed, (uint64_t)c * 16 + 2 * j);
        double u2 = laph::hash_u01(seed, (uint64_t)c * 16 + 2 * j + 1);
        xi[j] = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
      }
      // H = sum_a xi_a lambda_a / 2 (Gell-Mann basis)
      const double r3 = 1.0 / std::sqrt(3.0);
      cd H[9];
      H[0] = 0.5 * (xi[2] + xi[7] * r3);
      H[4] = 0.5 * (-xi[2] + xi[7] * r3);
      H[8] = 0.5 * (-2.0 * xi[7] * r3);
      H[1] = 0.5 * cd(xi[0], -xi[1]);
      H[3] = std::conj(H[1]);
      H[2] = 0.5 * cd(xi[3], -xi[4]);
      H[6] = std::conj(H[2]);
      H[5] = 0.5 * cd(xi[5], -xi[6]);
      H[7] = std::conj(H[5]);
      double w[3];
      cd Q[9];
      laph::herm3_eig(H, w, Q);
      cd e[3];
      for (int k = 0; k < 3; ++k) e[k] = std::exp(cd(0.0, eps * w[k]));
      int64_t site = c / 4;
      int mu = (int)(c % 4);
      int slot = (mu == 3) ? 0 : mu + 1;  // io.cpp:116-121
      double* dst = config + site * 72 + slot * 18;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          cd u = Q[3 * a] * e[0] * std::conj(Q[3 * b]) + Q[3 * a + 1] * e[1] * std::conj(Q[3 * b + 1]) +
                 Q[3 * a + 2] * e[2] * std::conj(Q[3 * b + 2]);
          dst[6 * a + 2 * b] = u.real();
          dst[6 * a + 2 * b + 1] = u.imag();
        }
    });
    return LAPH_OK;
  });
}

// ---------------------------------------------------------------------------
// LIME: record header = magic 0x456789ab (u32 BE), version (u16 BE), flags
// (u16 BE, 0x8000 message begin, 0x4000 message end), data length (u64 BE),
// 128-byte NUL-padded type; payload padded to a multiple of 8 bytes.
static uint64_t rd_be(const unsigned char* p, int nb) {
  uint64_t v = 0;
  for (int i = 0; i < nb; ++i) v = (v << 8) | p[i];
  return v;
}
static void wr_be(unsigned char* p, uint64_t v, int nb) {
  for (int i = nb - 1; i >= 0; --i) {
    p[i] = (unsigned char)(v & 0xff);
    v >>= 8;
  }
}
static double be_double(const unsigned char* p) {
  uint64_t v = rd_be(p, 8);
  double d;
  std::memcpy(&d, &v, 8);
  return d;
}

int laph_read_lime(const char* path, int T, int L, int ti, int tf, double* config) {
  LAPH_CHECK(path && config && T > 0 && L > 0 && ti >= 0 && tf >= ti && tf < T, LAPH_ERR_INVALID,
             "laph_read_lime: bad arguments (T=%d L=%d ti=%d tf=%d)", T, L, ti, tf);
  return laph::guarded([&]() -> int {
    FILE* f = std::fopen(path, "rb");
    LAPH_CHECK(f, LAPH_ERR_IO, "could not open %s", path);
    unsigned char hdr[144];
    uint64_t len = 0;
    bool found = false;
    while (std::fread(hdr, 1, 144, f) == 144) {
      if (rd_be(hdr, 4) != 0x456789abu) break;
      len = rd_be(hdr + 8, 8);
      char type[129];
      std::memcpy(type, hdr + 16, 128);
      type[128] = 0;
      if (std::strcmp(type, "ildg-binary-data") == 0) {
        found = true;
        break;
      }
      if (fseeko(f, (off_t)((len + 7) & ~7ull), SEEK_CUR) != 0) break;
    }
    if (!found) {
      std::fclose(f);
      return laph::set_error(LAPH_ERR_FORMAT, "no ildg-binary-data record found in %s", path);
    }
    const uint64_t V3 = (uint64_t)L * L * L;
    if (len != V3 * T * 72 * 8) {
      std::fclose(f);
      return laph::set_error(LAPH_ERR_FORMAT,
                             "wrong lattice size or precision in %s: %llu bytes, expected %llu",
                             path, (unsigned long long)len, (unsigned long long)(V3 * T * 576));
    }
    // skip to timeslice ti, then read whole timeslices (file order t,z,y,x)
    if (fseeko(f, (off_t)(V3 * 576 * (uint64_t)ti), SEEK_CUR) != 0) {
      std::fclose(f);
      return laph::set_error(LAPH_ERR_IO, "seek failed in %s", path);
    }
    std::vector<unsigned char> buf(V3 * 576);
    for (int t = ti; t <= tf; ++t) {
      if (std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
        std::fclose(f);
        return laph::set_error(LAPH_ERR_IO, "short read in %s at timeslice %d", path, t);
      }
      double* dst = config + (uint64_t)(t - ti) * V3 * 72;
      laph::parallel_for((int64_t)V3, [&](int64_t fi) {
        // fi = (z*L + y)*L + x in file order
        int64_t x = fi % L, y = (fi / L) % L, z = fi / ((int64_t)L * L);
        const unsigned char* src = buf.data() + fi * 576;
        double* p = dst + ((x * L + y) * L + z) * 72;
        int k = 0;
        for (int mu = 0; mu < 4; ++mu) {  // ILDG x,y,z,t -> slot 1,2,3,0
          double* q = p + (mu == 3 ? 0 : (mu + 1) * 18);
          for (int e = 0; e < 9; ++e, ++k) {
            q[2 * e] = be_double(src + 16 * k);
            q[2 * e + 1] = be_double(src + 16 * k + 8);
          }
        }
      });
    }
    std::fclose(f);
    return LAPH_OK;
  });
}

int laph_write_lime(const char* path, int T, int L, const double* config) {
  LAPH_CHECK(path && config && T > 0 && L > 0, LAPH_ERR_INVALID, "laph_write_lime: bad arguments");
  return laph::guarded([&]() -> int {
    FILE* f = std::fopen(path, "wb");
    LAPH_CHECK(f, LAPH_ERR_IO, "could not create %s", path);
    const uint64_t V3 = (uint64_t)L * L * L;
    auto header = [&](const char* type, uint64_t nbytes, bool mb, bool me) {
      unsigned char h[144] = {0};
      wr_be(h, 0x456789ab, 4);
      wr_be(h + 4, 1, 2);
      wr_be(h + 6, (mb ? 0x8000 : 0) | (me ? 0x4000 : 0), 2);
      wr_be(h + 8, nbytes, 8);
      std::strncpy((char*)h + 16, type, 127);
      return std::fwrite(h, 1, 144, f) == 144;
    };
    char xml[512];
    int nx = std::snprintf(xml, sizeof xml,
                           "<?xml version=\"1.0\" encoding=\"UTF-8\"?><ildgFormat><version>1.0"
                           "</version><field>su3gauge</field><precision>64</precision><lx>%d</lx>"
                           "<ly>%d</ly><lz>%d</lz><lt>%d</lt></ildgFormat>",
                           L, L, L, T);
    static const unsigned char zeros[8] = {0};
    bool ok = header("ildg-format", (uint64_t)nx, true, false) &&
              std::fwrite(xml, 1, nx, f) == (size_t)nx &&
              std::fwrite(zeros, 1, (8 - nx % 8) % 8, f) == (size_t)((8 - nx % 8) % 8) &&
              header("ildg-binary-data", V3 * T * 576, false, true);
    std::vector<unsigned char> buf(V3 * 576);
    for (int t = 0; ok && t < T; ++t) {
      const double* src = config + (uint64_t)t * V3 * 72;
      laph::parallel_for((int64_t)V3, [&](int64_t fi) {
        int64_t x = fi % L, y = (fi / L) % L, z = fi / ((int64_t)L * L);
        const double* p = src + ((x * L + y) * L + z) * 72;
        unsigned char* dst = buf.data() + fi * 576;
        int k = 0;
        for (int mu = 0; mu < 4; ++mu) {
          const double* q = p + (mu == 3 ? 0 : (mu + 1) * 18);
          for (int e = 0; e < 18; ++e, ++k) {
            uint64_t v;
            std::memcpy(&v, q + e, 8);
            wr_be(dst + 8 * k, v, 8);
          }
        }
      });
      ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    LAPH_CHECK(ok, LAPH_ERR_IO, "write failed for %s", path);
    return LAPH_OK;
  });
}

// ---------------------------------------------------------------------------
int laph_recover_spectrum(int nev, const double* theta, int deg, double lambda_l, double lambda_c,
                          double* evals) {
  LAPH_CHECK(nev >= 0 && deg >= 1 && theta && evals, LAPH_ERR_INVALID,
             "laph_recover_spectrum: bad arguments");
  for (int i = 0; i < nev; ++i) {
    double b = std::cosh(std::acosh(theta[i]) / deg);         // recover_spec.cpp:24-31
    evals[i] = (lambda_l - lambda_c) * (b - 1) * 0.5 - lambda_c;  // recover_spec.cpp:13-20
  }
  return LAPH_OK;
}

int laph_fix_phase(int n, int nev, const double* V, double* V_fix, double* phase) {
  LAPH_CHECK(n > 0 && nev >= 0 && V && V_fix && phase, LAPH_ERR_INVALID,
             "laph_fix_phase: bad arguments");
  const cd* v = reinterpret_cast<const cd*>(V);
  cd* vf = reinterpret_cast<cd*>(V_fix);
  for (int c = 0; c < nev; ++c) {
    double ph = std::arg(v[(size_t)c * n]);
    phase[c] = ph;
    cd fac = std::exp(cd(0.0, -ph));
    for (int r = 0; r < n; ++r) vf[(size_t)c * n + r] = fac * v[(size_t)c * n + r];
  }
  return LAPH_OK;
}

int laph_reset_phase(int n, int nev, const double* V, const double* want, double* V_reset) {
  LAPH_CHECK(n > 0 && nev >= 0 && V && want && V_reset, LAPH_ERR_INVALID,
             "laph_reset_phase: bad arguments");
  const cd* v = reinterpret_cast<const cd*>(V);
  cd* vr = reinterpret_cast<cd*>(V_reset);
  for (int c = 0; c < nev; ++c) {
    cd fac = std::exp(cd(0.0, -(std::arg(v[(size_t)c * n]) - want[c])));
    for (int r = 0; r < n; ++r) vr[(size_t)c * n + r] = fac * v[(size_t)c * n + r];
  }
  return LAPH_OK;
}

int laph_check_trace(int n, int nev, const double* V, double* trace, double* sum) {
  LAPH_CHECK(n > 0 && nev >= 0 && V, LAPH_ERR_INVALID, "laph_check_trace: bad arguments");
  const cd* v = reinterpret_cast<const cd*>(V);
  std::vector<cd> rows(nev);
  laph::parallel_for(nev, [&](int64_t i) {
    cd s_i = 0;
    for (int j = 0; j < nev; ++j) {
      cd acc = 0;
      const cd* a = v + (size_t)i * n;
      const cd* b = v + (size_t)j * n;
      for (int r = 0; r < n; ++r) acc += std::conj(a[r]) * b[r];
      s_i += acc;
    }
    rows[i] = s_i;
  });
  cd tr = 0, sm = 0;
  for (int i = 0; i < nev; ++i) {
    cd acc = 0;
    const cd* a = v + (size_t)i * n;
    for (int r = 0; r < n; ++r) acc += std::conj(a[r]) * a[r];
    tr += acc;
    sm += rows[i];
  }
  if (trace) { trace[0] = tr.real(); trace[1] = tr.imag(); }
  if (sum) { sum[0] = sm.real(); sum[1] = sm.imag(); }
  const double eps = 10e-10;  // read_write.cpp:24 (one-sided, as there)
  if (std::abs(tr.real()) - nev > eps || std::abs(tr.imag()) > eps ||
      std::abs(sm.real()) - nev > eps || std::abs(sm.imag()) > eps)
    return laph::set_error(LAPH_ERR_TRACE, "tr(V^H V) = (%.12g, %.3g), sum = (%.12g, %.3g), nev = %d",
                           tr.real(), tr.imag(), sm.real(), sm.imag(), nev);
  return LAPH_OK;
}

int laph_write_eig_sys_bin(const char* dir, const char* prefix, int config, int t, int n, int nev,
                           const double* V) {
  LAPH_CHECK(dir && prefix && V && n > 0 && nev > 0, LAPH_ERR_INVALID,
             "laph_write_eig_sys_bin: bad arguments");
  LAPH_TRY(laph_check_trace(n, nev, V, nullptr, nullptr));  // read_write.cpp:183-186
  char file[4096];
  std::snprintf(file, sizeof file, "%s/%s.%04d.%03d", dir, prefix, config, t);
  std::ofstream out(file, std::ofstream::binary);
  LAPH_CHECK(out, LAPH_ERR_IO, "could not create %s", file);
  out.write(reinterpret_cast<const char*>(V), (std::streamsize)2 * n * nev * sizeof(double));
  LAPH_CHECK(out.good(), LAPH_ERR_IO, "write incomplete: %s", file);
  return LAPH_OK;
}

int laph_write_eigenvalues_bin(const char* dir, const char* prefix, int config, int t, int nev,
                               const double* values) {
  LAPH_CHECK(dir && prefix && values && nev >= 0, LAPH_ERR_INVALID,
             "laph_write_eigenvalues_bin: bad arguments");
  char file[4096];
  std::snprintf(file, sizeof file, "%s/%s.%04d.%03d", dir, prefix, config, t);
  std::ofstream out(file, std::ofstream::binary);
  LAPH_CHECK(out, LAPH_ERR_IO, "could not create %s", file);
  out.write(reinterpret_cast<const char*>(values), (std::streamsize)nev * sizeof(double));
  LAPH_CHECK(out.good(), LAPH_ERR_IO, "write incomplete: %s", file);
  return LAPH_OK;
}

int laph_read_evectors_bin_ts(const char* dir, const char* prefix, int config, int t, int n, int nev,
                              double* V) {
  LAPH_CHECK(dir && prefix && V && n > 0 && nev > 0, LAPH_ERR_INVALID, "laph_read_evectors_bin_ts: bad arguments");
  char file[4096];
  std::snprintf(file, sizeof file, "%s/%s.%04d.%03d", dir, prefix, config, t);
  std::ifstream in(file, std::ifstream::binary);
  LAPH_CHECK(in, LAPH_ERR_IO, "could not open %s", file);
  for (int i = 0; i < nev; ++i) {  // one eigenvector after the other, as read_write.cpp:58-62
    in.read(reinterpret_cast<char*>(V + (size_t)2 * n * i), (std::streamsize)2 * n * sizeof(double));
    LAPH_CHECK(in.gcount() == (std::streamsize)(2 * n * sizeof(double)), LAPH_ERR_FORMAT,
               "%s: %d of %d eigenvectors of length %d in the file", file, i, nev, n);
  }
  return laph_check_trace(n, nev, V, nullptr, nullptr);  // read_write.cpp:63-66
}

int laph_read_eigenvalues_bin(const char* dir, const char* prefix, int config, int t, int nev, double* values) {
  LAPH_CHECK(dir && prefix && values && nev >= 0, LAPH_ERR_INVALID, "laph_read_eigenvalues_bin: bad arguments");
  char file[4096];
  std::snprintf(file, sizeof file, "%s/%s.%04d.%03d", dir, prefix, config, t);
  std::ifstream in(file, std::ifstream::binary);
  LAPH_CHECK(in, LAPH_ERR_IO, "could not open %s", file);
  in.read(reinterpret_cast<char*>(values), (std::streamsize)nev * sizeof(double));
  LAPH_CHECK(in.gcount() == (std::streamsize)(nev * sizeof(double)), LAPH_ERR_FORMAT, "%s: fewer than %d values",
             file, nev);
  return LAPH_OK;
}

int laph_read_eigenvalues_ascii(const char* prefix, int config, int t, int nev, double* values) {
  LAPH_CHECK(prefix && values && nev >= 0, LAPH_ERR_INVALID, "laph_read_eigenvalues_ascii: bad arguments");
  char file[4096];
  std::snprintf(file, sizeof file, "%s.%03d.%04d", prefix, t, config);  // read_write.cpp:104
  std::ifstream in(file);
  LAPH_CHECK(in, LAPH_ERR_IO, "could not open %s", file);
  for (int i = 0; i < nev; ++i) {
    in >> values[i];
    LAPH_CHECK(!in.fail(), LAPH_ERR_FORMAT, "%s: value %d of %d unreadable", file, i, nev);
  }
  return LAPH_OK;
}

}  // extern "C"
