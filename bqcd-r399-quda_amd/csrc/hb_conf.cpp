// hb_conf.cpp - BQCD configuration files, read and written on the host:
//
//  - one file per time slice t, written by field_io (comm/field_io_mpi.F90:
//    32-190) from conf_seq (conf.F90:219-260): for z, y, x (x fastest of the
//    three), for mu, the first two columns of U_mu(x) (row fastest), complex
//    float64, big-endian (swap_endian8, misc.F90:197-235);
//  - the third column is rebuilt for SU(3) as u_complete does (su3.F90:301-310);
//  - a ".info" text file: the ConfInfoHeader block (conf_info.F90:29-105) and
//    a CheckSum block with "file crc bytes" per slice (conf_info.F90:167-209),
//    the crc being POSIX cksum of the file (cksum.c).
//
// File names follow files.F90: <prog>.<run:3>.<ensemble:1>.<former:1>.<traj:5>
// with ".<t:2>.u" per slice and ".info" for the header.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "hb_internal.h"

namespace hb {
namespace {

// POSIX.2 cksum: CRC with generator 0x04C11DB7 over the bytes, then over the
// byte count (least significant byte first, only as many bytes as it has),
// then complemented.
struct Crc {
  uint32_t tab[256];
  Crc() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i << 24;
      for (int k = 0; k < 8; k++) c = (c & 0x80000000u) ? (c << 1) ^ 0x04C11DB7u : (c << 1);
      tab[i] = c;
    }
  }
};
const Crc &crc_table() {
  static const Crc t;
  return t;
}

struct Cksum {
  uint32_t crc = 0;
  uint64_t bytes = 0;
  void add(const unsigned char *p, size_t n) {
    const Crc &t = crc_table();
    for (size_t i = 0; i < n; i++) crc = (crc << 8) ^ t.tab[((crc >> 24) ^ p[i]) & 0xFF];
    bytes += n;
  }
  uint32_t get() const {
    const Crc &t = crc_table();
    uint32_t c = crc;
    for (uint64_t n = bytes; n > 0; n >>= 8) c = (c << 8) ^ t.tab[((c >> 24) ^ (n & 0xFF)) & 0xFF];
    return ~c;
  }
};

inline void put_be(unsigned char *p, double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  for (int k = 7; k >= 0; k--) {
    p[k] = (unsigned char)(u & 0xFF);
    u >>= 8;
  }
}
inline double get_be(const unsigned char *p) {
  uint64_t u = 0;
  for (int k = 0; k < 8; k++) u = (u << 8) | p[k];
  double v;
  std::memcpy(&v, &u, 8);
  return v;
}

// host gauge field in BQCD layout: complex u(row, col, site, eo, mu) with
// site_stride sites per (eo, mu) slab
struct HostGauge {
  double *u;
  int stride;
  double *at(int mu, int eo, int i, int row, int col) const {
    return u + 2 * (row + 3 * (col + 3 * ((size_t)i + (size_t)stride * (eo + 2 * (size_t)mu))));
  }
};

std::string dir_of(const std::string &path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

size_t slice_bytes(const int L[4]) { return (size_t)L[0] * L[1] * L[2] * 4 * 6 * 16; }

int write_slice(const int L[4], const HostGauge &g, int t, const std::string &path, uint32_t *crc,
                uint64_t *nbytes) {
  std::vector<unsigned char> buf(slice_bytes(L));
  unsigned char *p = buf.data();
  const int nxh = L[0] / 2;
  for (int z = 0; z < L[2]; z++)
    for (int y = 0; y < L[1]; y++)
      for (int x = 0; x < L[0]; x++) {
        const int eo = (x + y + z + t) & 1;
        const int i = x / 2 + nxh * (y + L[1] * (z + L[2] * t));
        for (int mu = 0; mu < 4; mu++)
          for (int col = 0; col < 2; col++)
            for (int row = 0; row < 3; row++) {
              const double *v = g.at(mu, eo, i, row, col);
              put_be(p, v[0]);
              put_be(p + 8, v[1]);
              p += 16;
            }
      }
  FILE *f = fopen(path.c_str(), "wb");
  HB_CHECK(f, HB_E_ARG, "conf: cannot write %s: %s", path.c_str(), strerror(errno));
  const size_t w = fwrite(buf.data(), 1, buf.size(), f);
  const int cl = fclose(f);
  HB_CHECK(w == buf.size() && cl == 0, HB_E_ARG, "conf: short write to %s", path.c_str());
  Cksum ck;
  ck.add(buf.data(), buf.size());
  *crc = ck.get();
  *nbytes = ck.bytes;
  return HB_OK;
}

int read_slice(const int L[4], const HostGauge &g, int t, const std::string &path, uint32_t crc_expect,
               uint64_t bytes_expect) {
  const size_t n = slice_bytes(L);
  std::vector<unsigned char> buf(n);
  FILE *f = fopen(path.c_str(), "rb");
  HB_CHECK(f, HB_E_ARG, "conf: cannot open %s: %s", path.c_str(), strerror(errno));
  const size_t r = fread(buf.data(), 1, n, f);
  const int extra = fgetc(f);
  fclose(f);
  HB_CHECK(r == n && extra == EOF, HB_E_ARG, "conf: %s has the wrong size for a %dx%dx%d time slice",
           path.c_str(), L[0], L[1], L[2]);
  Cksum ck;
  ck.add(buf.data(), n);
  HB_CHECK(ck.bytes == bytes_expect && ck.get() == crc_expect, HB_E_ARG,
           "conf: checksum of %s is %u %llu, the info file says %u %llu", path.c_str(), ck.get(),
           (unsigned long long)ck.bytes, crc_expect, (unsigned long long)bytes_expect);
  const unsigned char *p = buf.data();
  const int nxh = L[0] / 2;
  for (int z = 0; z < L[2]; z++)
    for (int y = 0; y < L[1]; y++)
      for (int x = 0; x < L[0]; x++) {
        const int eo = (x + y + z + t) & 1;
        const int i = x / 2 + nxh * (y + L[1] * (z + L[2] * t));
        for (int mu = 0; mu < 4; mu++) {
          for (int col = 0; col < 2; col++)
            for (int row = 0; row < 3; row++) {
              double *v = g.at(mu, eo, i, row, col);
              v[0] = get_be(p);
              v[1] = get_be(p + 8);
              p += 16;
            }
          // third column: conj of the cross product of the first two (su3.F90:301-310)
          auto c = [&](int row, int col, double &re, double &im) {
            const double *v = g.at(mu, eo, i, row, col);
            re = v[0];
            im = v[1];
          };
          for (int row = 0; row < 3; row++) {
            const int r1 = (row + 1) % 3, r2 = (row + 2) % 3;
            double ar, ai, br, bi, cr, ci, dr, di;
            c(r1, 0, ar, ai);  // u(r1,1)
            c(r2, 1, br, bi);  // u(r2,2)
            c(r2, 0, cr, ci);  // u(r2,1)
            c(r1, 1, dr, di);  // u(r1,2)
            const double re = (ar * br - ai * bi) - (cr * dr - ci * di);
            const double im = (ar * bi + ai * br) - (cr * di + ci * dr);
            double *v = g.at(mu, eo, i, row, 2);
            v[0] = re;
            v[1] = -im;
          }
        }
      }
  return HB_OK;
}

}  // namespace
}  // namespace hb

using namespace hb;

extern "C" {

int hb_conf_write(const int L[4], const int bc[4], const double *u, int site_stride, const char *dir,
                  int run, int traj, int ensemble, int former, double beta, double kappa, double csw,
                  double h, double plaq_energy, char *info_path, int info_path_len) {
  HB_CHECK(L && bc && u && dir, HB_E_ARG, "hb_conf_write: null argument");
  for (int m = 0; m < 4; m++)
    HB_CHECK(L[m] >= 2 && L[m] % 2 == 0, HB_E_ARG, "hb_conf_write: L[%d] = %d", m, L[m]);
  const int vh = L[0] * L[1] * L[2] * L[3] / 2;
  HB_CHECK(site_stride >= vh, HB_E_ARG, "hb_conf_write: site_stride < volh");
  HB_CHECK(run >= 0 && run <= 999 && traj >= 0 && traj <= 99999 && ensemble >= 1 && ensemble <= 9 &&
               former >= 1 && former <= 9 && L[3] <= 100,
           HB_E_ARG, "hb_conf_write: run, trajectory or ensemble outside the file-name formats");
  const HostGauge g{const_cast<double *>(u), site_stride};
  char base[256];
  snprintf(base, sizeof base, "bqcd.%03d.%d.%d.%05d", run, ensemble, former, traj);
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  std::vector<uint32_t> crc(L[3]);
  std::vector<uint64_t> nb(L[3]);
  std::vector<std::string> names(L[3]);
  for (int t = 0; t < L[3]; t++) {
    char name[300];
    snprintf(name, sizeof name, "%s.%02d.u", base, t);
    names[t] = name;
    HB_TRY(write_slice(L, g, t, d + name, &crc[t], &nb[t]));
  }
  const std::string info = d + base + ".info";
  FILE *f = fopen(info.c_str(), "w");
  HB_CHECK(f, HB_E_ARG, "hb_conf_write: cannot write %s: %s", info.c_str(), strerror(errno));
  char date[64];
  const time_t now = time(nullptr);
  strftime(date, sizeof date, "%Y-%m-%d %H:%M:%S", localtime(&now));
  char host[64] = "unknown";
  if (FILE *hn = fopen("/proc/sys/kernel/hostname", "r")) {
    if (fgets(host, sizeof host, hn)) host[strcspn(host, "\n")] = 0;
    fclose(hn);
  }
  const double par[5] = {beta, kappa, csw, csw * kappa, h};
  const char *key[5] = {"beta", "kappa", "csw", "csw_kappa", "h"};
  fprintf(f, " >BeginConfInfoHeader\n Format 3\n Program bqcd-r399-quda_amd hip\n Run %d\n Traj %d\n", run,
          traj);
  fprintf(f, " Host %s\n Date %s\n L %3d%3d%3d%3d\n bc_fermions %3d%3d%3d%3d\n REAL_kind 8\n", host, date,
          L[0], L[1], L[2], L[3], bc[0], bc[1], bc[2], bc[3]);
  fprintf(f, " PlaqEnergy %25.14E\n", plaq_energy);
  for (int e = 0; e < 2; e++) {
    fprintf(f, " %sensemble %d\n", e ? "former_" : "", e ? former : ensemble);
    for (int k = 0; k < 5; k++) fprintf(f, " %s%s %.15g\n", e ? "former_" : "", key[k], par[k]);
  }
  fprintf(f, " >EndConfInfoHeader\n >BeginCheckSum\n");
  for (int t = 0; t < L[3]; t++)
    fprintf(f, " %s %u %llu\n", names[t].c_str(), crc[t], (unsigned long long)nb[t]);
  fprintf(f, " >EndCheckSum\n");
  const int cl = fclose(f);
  HB_CHECK(cl == 0, HB_E_ARG, "hb_conf_write: cannot close %s", info.c_str());
  if (info_path && info_path_len > 0) snprintf(info_path, info_path_len, "%s", info.c_str());
  return HB_OK;
}

int hb_conf_read(const char *info_file, const int L[4], double *u, int site_stride, double *plaq_energy) {
  HB_CHECK(info_file && L && u, HB_E_ARG, "hb_conf_read: null argument");
  const int vh = L[0] * L[1] * L[2] * L[3] / 2;
  HB_CHECK(site_stride >= vh, HB_E_ARG, "hb_conf_read: site_stride < volh");
  FILE *f = fopen(info_file, "r");
  HB_CHECK(f, HB_E_ARG, "hb_conf_read: cannot open %s: %s", info_file, strerror(errno));
  std::vector<std::string> lines;
  char buf[1024];
  while (fgets(buf, sizeof buf, f)) lines.push_back(trim(buf));
  fclose(f);
  int format = -1, rkind = -1, fl[4] = {0, 0, 0, 0};
  double plaq = NAN;
  std::vector<std::string> names;
  std::vector<uint32_t> crc;
  std::vector<uint64_t> nb;
  bool in_ck = false;
  for (const std::string &l : lines) {
    if (l == ">BeginCheckSum") { in_ck = true; continue; }
    if (l == ">EndCheckSum") { in_ck = false; continue; }
    if (in_ck) {
      char name[512];
      unsigned long long c = 0, b = 0;
      HB_CHECK(sscanf(l.c_str(), "%511s %llu %llu", name, &c, &b) == 3, HB_E_ARG,
               "hb_conf_read: bad checksum line '%s'", l.c_str());
      names.push_back(name);
      crc.push_back((uint32_t)c);
      nb.push_back(b);
      continue;
    }
    if (sscanf(l.c_str(), "Format %d", &format) == 1) continue;
    if (sscanf(l.c_str(), "REAL_kind %d", &rkind) == 1) continue;
    if (sscanf(l.c_str(), "L %d %d %d %d", &fl[0], &fl[1], &fl[2], &fl[3]) == 4) continue;
    if (sscanf(l.c_str(), "PlaqEnergy %lf", &plaq) == 1) continue;
  }
  HB_CHECK(format == 3, HB_E_ARG, "hb_conf_read: %s: info format %d, expected 3", info_file, format);
  HB_CHECK(rkind == 8, HB_E_ARG, "hb_conf_read: %s: REAL_kind %d, expected 8", info_file, rkind);
  for (int m = 0; m < 4; m++)
    HB_CHECK(fl[m] == L[m], HB_E_ARG, "hb_conf_read: %s holds a %dx%dx%dx%d lattice", info_file, fl[0], fl[1],
             fl[2], fl[3]);
  HB_CHECK((int)names.size() == L[3], HB_E_ARG, "hb_conf_read: %s lists %d time slices, expected %d",
           info_file, (int)names.size(), L[3]);
  const std::string d = dir_of(info_file);
  const HostGauge g{u, site_stride};
  for (int t = 0; t < L[3]; t++) {
    const std::string path = names[t][0] == '/' ? names[t] : d + names[t];
    HB_TRY(read_slice(L, g, t, path, crc[t], nb[t]));
  }
  if (plaq_energy) *plaq_energy = plaq;
  return HB_OK;
}

}  // extern "C"
