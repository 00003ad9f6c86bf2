// cc_io.cpp -- the reference's formatted point files (solution files
// 3Dnnn.dat, fluc.dat, qInvariant.dat), written and read in parallel on the
// host.  At nP = [256,128,192] one solution file is 1.2 GB of text: the
// reference writes it record by record through the Fortran runtime
// (main_v1.9.f90:759-773, format 101 = (3(1X,F10.6),7(1X,E22.15))) and
// reads it the same way (:222-251).  Here every thread formats or parses a
// contiguous range of records; the output is byte-identical to the Fortran
// runtime's.
//
// Fortran edit descriptors reproduced (flang's runtime conventions):
//   Fw.d   : C's %w.df (correctly rounded, '-' kept on negative zero and on
//            negatives that round to zero); '*' x w when it does not fit;
//   E22.15 : [-]0.ddddddddddddddd E+ee from the correctly rounded 15
//            significant digits; a 3-digit exponent drops the 'E'
//            (0.250000000000000-299); zero prints E+00;
//   NaN / Inf / -Inf right-justified in the field.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cc_abi.h"

namespace {

thread_local std::string g_ioerr;

int io_fail(int code, const std::string& msg) {
  g_ioerr = msg;
  return code;
}

int nthreads_of(int want) {
  if (want > 0) return want;
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h, 16u));
}

// Fw.d into p[0..w); returns w.
int put_F(char* p, double x, int w, int d) {
  char b[64];
  int n;
  if (std::isnan(x)) n = snprintf(b, sizeof b, "NaN");
  else if (std::isinf(x)) n = snprintf(b, sizeof b, x > 0 ? "Inf" : "-Inf");
  else n = snprintf(b, sizeof b, "%.*f", d, x);
  if (n > w || n < 0) {
    memset(p, '*', w);
  } else {
    memset(p, ' ', w - n);
    memcpy(p + w - n, b, n);
  }
  return w;
}

// The 15 correctly rounded significant digits of |x| (x finite, nonzero)
// and its decimal exponent e10 (|x| ~ d.ddd 10^e10), without printf: |x| is
// scaled by an exact power of ten in x87 extended precision (64-bit
// significand, one rounding, relative error <= 2^-64, i.e. an absolute error
// below 1e-4 on the scaled value in [1e14, 1e15)), then rounded to an
// integer.  Returns false (the caller uses snprintf) when the scaled value
// lies within 1e-3 of a rounding tie or the power of ten is not exact in
// extended precision (|14 - e10| > 27).
bool digits15(double x, uint64_t& m, int& e10) {
  static const long double p10[28] = {
      1e0L,  1e1L,  1e2L,  1e3L,  1e4L,  1e5L,  1e6L,  1e7L,  1e8L,  1e9L,
      1e10L, 1e11L, 1e12L, 1e13L, 1e14L, 1e15L, 1e16L, 1e17L, 1e18L, 1e19L,
      1e20L, 1e21L, 1e22L, 1e23L, 1e24L, 1e25L, 1e26L, 1e27L};
  const long double ax = std::fabs((long double)x);
  // decimal exponent from the binary one (within 1; corrected below)
  int e = (int)std::floor(std::ilogb(x) * 0.30102999566398120);
  for (int pass = 0; pass < 3; pass++) {
    const int k = 14 - e;
    if (k > 27 || k < -27) return false;
    const long double sc = k >= 0 ? ax * p10[k] : ax / p10[-k];
    if (sc < 1e14L) { e -= 1; continue; }
    if (sc >= 1e15L) { e += 1; continue; }
    const long double fl = std::floor(sc);
    const long double fr = sc - fl;
    if (fr > 0.499L && fr < 0.501L) return false;
    m = (uint64_t)fl + (fr > 0.5L ? 1 : 0);
    if (m == 1000000000000000ull) {
      m = 100000000000000ull;
      e += 1;
    }
    e10 = e;
    return true;
  }
  return false;
}

// E22.15 into p[0..22); returns 22.
int put_E22_15(char* p, double x) {
  const int w = 22;
  char b[64];
  if (x != 0.0 && std::isfinite(x)) {
    uint64_t m;
    int e10;
    if (digits15(x, m, e10)) {
      char* q = p;
      *q++ = std::signbit(x) ? '-' : ' ';
      *q++ = '0';
      *q++ = '.';
      for (int d = 14; d >= 0; d--) {
        q[d] = (char)('0' + m % 10);
        m /= 10;
      }
      q += 15;
      const int ex = e10 + 1, ae = ex < 0 ? -ex : ex;
      if (ae <= 99) {
        *q++ = 'E';
        *q++ = ex < 0 ? '-' : '+';
        *q++ = (char)('0' + ae / 10);
        *q++ = (char)('0' + ae % 10);
      } else {
        *q++ = ex < 0 ? '-' : '+';
        *q++ = (char)('0' + ae / 100);
        *q++ = (char)('0' + (ae / 10) % 10);
        *q++ = (char)('0' + ae % 10);
      }
      return w;
    }
  }
  if (std::isnan(x) || std::isinf(x)) {
    const char* s = std::isnan(x) ? "NaN" : (x > 0 ? "Inf" : "-Inf");
    const int n = (int)strlen(s);
    memset(p, ' ', w - n);
    memcpy(p + w - n, s, n);
    return w;
  }
  // d.dddddddddddddd E+xx : 15 significant digits, correctly rounded
  snprintf(b, sizeof b, "%.14E", x);
  const char* s = b;
  const bool neg = *s == '-';
  if (neg) s++;
  char dig[15];
  dig[0] = s[0];
  memcpy(dig + 1, s + 2, 14);
  const char* e = strchr(s, 'E');
  int ex = atoi(e + 1);
  if (x != 0.0) ex += 1;  // 0.d... form
  char* q = p;
  *q++ = neg ? '-' : ' ';
  *q++ = '0';
  *q++ = '.';
  memcpy(q, dig, 15);
  q += 15;
  const int ae = ex < 0 ? -ex : ex;
  if (ae <= 99) {
    *q++ = 'E';
    *q++ = ex < 0 ? '-' : '+';
    *q++ = (char)('0' + ae / 10);
    *q++ = (char)('0' + ae % 10);
  } else {
    *q++ = ex < 0 ? '-' : '+';
    *q++ = (char)('0' + ae / 100);
    *q++ = (char)('0' + (ae / 10) % 10);
    *q++ = (char)('0' + ae % 10);
  }
  return w;
}

// Parse one number token [s, e): fixed-point or E/D form, with the Fortran
// exponent without a letter (0.25-299).  Exact fast path for <= 15
// significant digits and |power of ten| <= 22 (both operands exact doubles,
// one rounding); strtod otherwise.
double parse_num(const char* s, const char* e) {
  const char* p = s;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  uint64_t m = 0;
  int nd = 0, scale = 0;
  bool any = false, ok = true;
  while (p < e && *p >= '0' && *p <= '9') {
    any = true;
    if (m || *p != '0') {
      if (nd < 19) { m = m * 10 + (uint64_t)(*p - '0'); nd++; } else { scale++; ok = false; }
    }
    p++;
  }
  if (p < e && *p == '.') {
    p++;
    while (p < e && *p >= '0' && *p <= '9') {
      any = true;
      if (m || *p != '0') {
        if (nd < 19) { m = m * 10 + (uint64_t)(*p - '0'); nd++; scale--; } else ok = false;
      } else {
        scale--;
      }
      p++;
    }
  }
  int ex = 0;
  if (p < e && (*p == 'E' || *p == 'e' || *p == 'D' || *p == 'd')) p++;
  if (p < e && (*p == '+' || *p == '-')) {
    const bool en = *p++ == '-';
    int v = 0;
    while (p < e && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
    ex = en ? -v : v;
  }
  if (!any || p != e) ok = false;
  if (ok && nd <= 15) {
    static const double p10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                   1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                   1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const int pw = scale + ex;
    double v = (double)m;
    if (m == 0) return neg ? -0.0 : 0.0;
    if (pw >= 0 && pw <= 22) return neg ? -(v * p10[pw]) : v * p10[pw];
    if (pw < 0 && pw >= -22) return neg ? -(v / p10[-pw]) : v / p10[-pw];
  }
  // general case (including NaN, Inf and a letterless exponent)
  std::string t(s, e);
  const size_t k = t.find_first_of("+-", 1);
  if (k != std::string::npos && t.find_first_of("EeDd") == std::string::npos &&
      t[k - 1] >= '0' && t[k - 1] <= '9')
    t.insert(k, "E");
  for (char& c : t)
    if (c == 'D' || c == 'd') c = 'E';
  return strtod(t.c_str(), nullptr);
}

bool is_blank(const char* s, const char* e) {
  for (; s < e; s++)
    if (*s != ' ' && *s != '\t' && *s != '\r') return false;
  return true;
}

}  // namespace

extern "C" {

const char* cc_io_last_error(void) { return g_ioerr.c_str(); }

int cc_write_points(const char* path, const char* header, int n1, int n2, int n3,
                    const double* x1, const double* x2, const double* x3, int nf,
                    const double* fields, const int* pattern, int blank_records, int nthreads) {
  if (!path || !header || !x1 || !x2 || !x3 || (nf > 0 && !fields) || n1 < 1 || n2 < 1 ||
      n3 < 1 || nf < 0)
    return io_fail(CC_EINVAL, "bad argument");
  const long long N = (long long)n1 * n2 * n3;
  const int rec = 3 * 11 + nf * 23 + (pattern ? 2 : 0) + 1;
  const int bl = blank_records ? 1 : 0;
  const long long plane_bytes = (long long)bl + (long long)n2 * (bl + (long long)n1 * rec);
  const size_t hlen = strlen(header);
  const long long total = (long long)hlen + plane_bytes * n3;
  const int fd = open(path, O_CREAT | O_TRUNC | O_WRONLY, 0644);
  if (fd < 0) return io_fail(CC_EINVAL, std::string("cannot open ") + path + ": " + strerror(errno));
  if (ftruncate(fd, total) != 0 || pwrite(fd, header, hlen, 0) != (ssize_t)hlen) {
    close(fd);
    return io_fail(CC_EINVAL, std::string("cannot write ") + path);
  }
  const int T = std::min(nthreads_of(nthreads), n3);
  std::vector<int> bad(T, 0);
  // the coordinate fields repeat across records: format each once
  std::vector<char> fx1((size_t)n1 * 11), fx2((size_t)n2 * 11), fx3((size_t)n3 * 11);
  auto prefmt = [](std::vector<char>& f, const double* x, int n) {
    for (int q = 0; q < n; q++) {
      f[(size_t)q * 11] = ' ';
      put_F(&f[(size_t)q * 11 + 1], x[q], 10, 6);
    }
  };
  prefmt(fx1, x1, n1);
  prefmt(fx2, x2, n2);
  prefmt(fx3, x3, n3);
  auto work = [&](int t) {
    const int k0 = (int)((long long)n3 * t / T), k1 = (int)((long long)n3 * (t + 1) / T);
    // format a bounded number of planes at a time
    const long long maxbuf = 64ll << 20;
    const int step = (int)std::max(1LL, std::min((long long)(k1 - k0), maxbuf / std::max(1LL, plane_bytes)));
    std::vector<char> buf((size_t)(plane_bytes * step));
    for (int kb = k0; kb < k1; kb += step) {
      const int ke = std::min(k1, kb + step);
      char* q = buf.data();
      for (int k = kb; k < ke; k++) {
        if (bl) *q++ = '\n';
        for (int j = 0; j < n2; j++) {
          if (bl) *q++ = '\n';
          for (int i = 0; i < n1; i++) {
            const long long n = ((long long)k * n2 + j) * n1 + i;
            memcpy(q, &fx1[(size_t)i * 11], 11);
            memcpy(q + 11, &fx2[(size_t)j * 11], 11);
            memcpy(q + 22, &fx3[(size_t)k * 11], 11);
            q += 33;
            for (int c = 0; c < nf; c++) {
              *q++ = ' ';
              q += put_E22_15(q, fields[(long long)c * N + n]);
            }
            if (pattern) {
              const int v = pattern[n];
              *q++ = ' ';
              *q++ = (v >= 0 && v <= 9) ? (char)('0' + v) : '*';
            }
            *q++ = '\n';
          }
        }
      }
      const long long off = (long long)hlen + plane_bytes * kb, len = q - buf.data();
      if (pwrite(fd, buf.data(), len, off) != len) bad[t] = 1;
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < T; t++) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  close(fd);
  for (int b : bad)
    if (b) return io_fail(CC_EINVAL, std::string("short write to ") + path);
  return 0;
}

int cc_read_points(const char* path, int nheader, int n1, int n2, int n3, int nf, int has_pattern,
                   double* x1, double* x2, double* x3, double* fields, int nthreads) {
  if (!path || n1 < 1 || n2 < 1 || n3 < 1 || nf < 0 || nheader < 0)
    return io_fail(CC_EINVAL, "bad argument");
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return io_fail(CC_EINVAL, std::string("cannot open ") + path + ": " + strerror(errno));
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size == 0) {
    close(fd);
    return io_fail(CC_EINVAL, std::string("empty file ") + path);
  }
  const size_t size = (size_t)st.st_size;
  void* mp = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (mp == MAP_FAILED) return io_fail(CC_EINVAL, std::string("cannot map ") + path);
  const char* base = (const char*)mp;
  const char* end = base + size;
  const char* body = base;
  for (int h = 0; h < nheader && body < end; h++) {
    const char* nl = (const char*)memchr(body, '\n', end - body);
    body = nl ? nl + 1 : end;
  }
  const long long N = (long long)n1 * n2 * n3;
  const int ntok = 3 + nf + (has_pattern ? 1 : 0);
  if (ntok > 64) {
    munmap(mp, size);
    return io_fail(CC_EINVAL, "at most 60 fields per record");
  }
  const int T = (int)std::max(1LL, std::min((long long)nthreads_of(nthreads), (long long)(end - body) / 4096 + 1));
  // chunk t owns the lines that start in [b_t, b_{t+1})
  std::vector<const char*> cut(T + 1);
  for (int t = 0; t <= T; t++) {
    const char* c = body + (size_t)((end - body) * (long double)t / T);
    if (t > 0 && t < T) {
      while (c < end && c[-1] != '\n') c++;
    }
    cut[t] = t == T ? end : c;
  }
  std::vector<long long> count(T, 0);
  auto lines = [&](int t, auto&& fn) {
    const char* s = cut[t];
    while (s < cut[t + 1]) {
      const char* nl = (const char*)memchr(s, '\n', end - s);
      const char* e = nl ? nl : end;
      if (!is_blank(s, e)) fn(s, e);
      s = e + 1;
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t] { lines(t, [&](const char*, const char*) { count[t]++; }); });
    for (auto& x : th) x.join();
  }
  long long tot = 0;
  std::vector<long long> first(T);
  for (int t = 0; t < T; t++) {
    first[t] = tot;
    tot += count[t];
  }
  if (tot != N) {
    munmap(mp, size);
    return io_fail(CC_EINVAL, std::string(path) + ": " + std::to_string(tot) + " records for a " +
                                  std::to_string(n1) + "x" + std::to_string(n2) + "x" +
                                  std::to_string(n3) + " grid");
  }
  std::vector<long long> badrec(T, -1);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t] {
        long long n = first[t];
        lines(t, [&](const char* s, const char* e) {
          double v[64];
          int k = 0;
          const char* p = s;
          while (p < e && k < ntok) {
            while (p < e && (*p == ' ' || *p == '\t')) p++;
            const char* q = p;
            while (q < e && *q != ' ' && *q != '\t' && *q != '\r') q++;
            if (q > p) v[k++] = parse_num(p, q);
            p = q;
          }
          if (k < ntok && badrec[t] < 0) badrec[t] = n;
          const int i = (int)(n % n1), j = (int)((n / n1) % n2), kk = (int)(n / ((long long)n1 * n2));
          if (x1 && j == 0 && kk == 0) x1[i] = v[0];
          if (x2 && i == 0 && kk == 0) x2[j] = v[1];
          if (x3 && i == 0 && j == 0) x3[kk] = v[2];
          if (fields)
            for (int c = 0; c < nf; c++) fields[(long long)c * N + n] = v[3 + c];
          n++;
        });
      });
    for (auto& x : th) x.join();
  }
  munmap(mp, size);
  for (long long b : badrec)
    if (b >= 0) return io_fail(CC_EINVAL, std::string(path) + ": short record " + std::to_string(b + 1));
  return 0;
}

int cc_write_solution(const char* path, double time, int nstep, int n1, int n2, int n3,
                      const double* x1, const double* x2, const double* x3, const double* var,
                      const int* pattern, int nthreads) {
  // header of main_v1.9.f90:759-761 (formats 103, list-directed, 104)
  char t[32], ns[32], d[3][8];
  put_F(t, time, 20, 10);
  t[20] = 0;
  snprintf(ns, sizeof ns, "%9d", nstep);
  if (strlen(ns) > 9) strcpy(ns, "*********");
  const int dims[3] = {n3, n2, n1};
  for (int q = 0; q < 3; q++) {
    snprintf(d[q], sizeof d[q], "%3d", dims[q]);
    if (strlen(d[q]) > 3) strcpy(d[q], "***");
  }
  std::string h = std::string("TITLE =\"") + t + ns + "\"\n";
  h += pattern ? " Variables =\"x\",\"y\",\"z\",\"Rho\",\"U\",\"V\",\"W\",\"E\",\"T\",\"P\",\"Pattern\"\n"
               : " Variables =\"x\",\"y\",\"z\",\"Rho\",\"U\",\"V\",\"W\",\"E\",\"T\",\"P\"\n";
  h += std::string("ZONE k=") + d[0] + ",j=" + d[1] + ",i=" + d[2] + ",DATAPACKING=\"POINT\"\n";
  return cc_write_points(path, h.c_str(), n1, n2, n3, x1, x2, x3, 7, var, pattern, 1, nthreads);
}

int cc_read_solution(const char* path, int n1, int n2, int n3, double* x1, double* x2,
                     double* x3, double* var, double* time, int* nstep, int nthreads) {
  if (!path) return io_fail(CC_EINVAL, "null path");
  FILE* f = fopen(path, "r");
  if (!f) return io_fail(CC_EINVAL, std::string("cannot open ") + path);
  char line[512] = {0}, vars[512] = {0};
  const bool ok = fgets(line, sizeof line, f) != nullptr;
  // the halo-corrected solver's files add an integer "Pattern" column after
  // the ten values (v1.10/codes/main_v1.9.f90:777,788)
  const bool has_pattern = fgets(vars, sizeof vars, f) != nullptr && strstr(vars, "Pattern");
  fclose(f);
  if (!ok || strlen(line) < 37) return io_fail(CC_EINVAL, std::string(path) + ": bad title line");
  // (8X,F20.10,I9,1X), main_v1.9.f90:227-231
  if (time) *time = parse_num(line + 8 + strspn(line + 8, " "), line + 28);
  if (nstep) {
    char b[10];
    memcpy(b, line + 28, 9);
    b[9] = 0;
    *nstep = atoi(b);
  }
  return cc_read_points(path, 3, n1, n2, n3, 7, has_pattern ? 1 : 0, x1, x2, x3, var, nthreads);
}

int cc_solution_dims(const char* path, int* n1, int* n2, int* n3) {
  if (!path || !n1 || !n2 || !n3) return io_fail(CC_EINVAL, "null argument");
  FILE* f = fopen(path, "r");
  if (!f) return io_fail(CC_EINVAL, std::string("cannot open ") + path);
  char line[512];
  int k = 0, j = 0, i = 0, got = 0;
  for (int l = 0; l < 3 && fgets(line, sizeof line, f); l++) {
    const char* z = strstr(line, "ZONE");
    if (z && sscanf(z, "ZONE k=%d ,j=%d ,i=%d", &k, &j, &i) == 3) got = 1;
  }
  fclose(f);
  if (!got) return io_fail(CC_EINVAL, std::string(path) + ": no ZONE k=,j=,i= line");
  *n1 = i;
  *n2 = j;
  *n3 = k;
  return 0;
}

int cc_format_field(double x, int kind, char* out) {
  // kind 0: E22.15, 1: F10.6; out receives the field and a NUL
  if (!out) return CC_EINVAL;
  const int w = kind == 0 ? put_E22_15(out, x) : put_F(out, x, 10, 6);
  out[w] = 0;
  return 0;
}

}  // extern "C"
