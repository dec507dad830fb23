// semx_io.cpp -- semtex field files (.fld / .rst / .chk) for device fields.
//
// A field file is what writeField (src/misc.cpp:106-180) writes and the
// Domain stream input (src/domain.cpp:329-398) reads: a 10-line header,
// each line a value left-justified in 25 columns and its label (Session,
// Created, "Nr, Ns, Nz, Elements", Step, Time, Time step, Kinvis, Beta,
// Fields written, Format), then for every field named on the "Fields
// written" line its nz physical-space planes of nPlane doubles (no plane
// padding) in the writer's byte order.  A run that dumps periodically
// appends one such block per dump to its .fld file (Domain::dump,
// src/domain.cpp:188-248).
//
// The fields handed in and out are the library's device fields in Fourier
// space (nz planes of planeSize doubles).  As Domain::dump does, a dump is
// transformed to physical space with the Nyquist plane zeroed first
// (Domain::transform, src/domain.cpp:250-264); a read is transformed
// forward and its Nyquist plane zeroed, as Domain::restart does
// (src/domain.cpp:152-180).  The transforms run on the device through
// semx_dft_r, on scratch copies; plain file I/O and byte order on the host.
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "semx.h"
#include "semx_internal.h"

namespace {

bool host_little_endian() {
  const uint16_t one = 1;
  unsigned char b;
  std::memcpy(&b, &one, 1);
  return b == 1;
}

void byte_reverse(double* x, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    unsigned char* c = reinterpret_cast<unsigned char*>(x + i);
    std::reverse(c, c + 8);
  }
}

// One Fourier-space field (device, nz*planeSize) to physical space on the
// host (nz*nPlane): Nyquist plane zeroed, inverse transform (nz = 2: plane
// 0 copied to plane 1, AuxField::transform, src/auxfield.cpp:1153-1158).
int to_physical(const semx::Mesh& M, int nz, const double* dev, double* scratch,
                std::vector<double>& host, std::vector<double>& out, void* stream) {
  const size_t ps = M.planeSize, n = (size_t)nz * ps;
  int st = semx_synchronize();
  if (!st) st = semx_memcpy_dtoh(host.data(), dev, n * sizeof(double));
  if (st) return st;
  if (nz > 1) std::fill(host.begin() + ps, host.begin() + 2 * ps, 0.0);
  if (nz == 2) {
    std::copy(host.begin(), host.begin() + ps, host.begin() + ps);
  } else if (nz > 2) {
    if ((st = semx_memcpy_htod(scratch, host.data(), n * sizeof(double)))) return st;
    if ((st = semx_dft_r(scratch, nz, (int)ps, SEMX_INVERSE, stream))) return st;
    if ((st = semx_synchronize())) return st;
    if ((st = semx_memcpy_dtoh(host.data(), scratch, n * sizeof(double)))) return st;
  }
  out.resize((size_t)nz * M.nPlane);
  for (int k = 0; k < nz; ++k)
    std::copy(host.begin() + (size_t)k * ps, host.begin() + (size_t)k * ps + M.nPlane,
              out.begin() + (size_t)k * M.nPlane);
  return SEMX_OK;
}

// Physical-space planes (host, nz*nPlane) to a Fourier-space device field:
// forward transform (nz = 2: plane 1 zeroed), then the Nyquist plane zeroed.
int to_fourier(const semx::Mesh& M, int nz, const double* phys, double* dev,
               std::vector<double>& host, void* stream) {
  const size_t ps = M.planeSize, n = (size_t)nz * ps;
  std::fill(host.begin(), host.end(), 0.0);
  for (int k = 0; k < nz; ++k)
    std::copy(phys + (size_t)k * M.nPlane, phys + (size_t)(k + 1) * M.nPlane,
              host.begin() + (size_t)k * ps);
  if (nz == 2) std::fill(host.begin() + ps, host.begin() + 2 * ps, 0.0);
  int st = semx_memcpy_htod(dev, host.data(), n * sizeof(double));
  if (st) return st;
  if (nz > 2) {
    if ((st = semx_dft_r(dev, nz, (int)ps, SEMX_FORWARD, stream))) return st;
    if ((st = semx_synchronize())) return st;
    if ((st = semx_memcpy_dtoh(host.data(), dev, n * sizeof(double)))) return st;
    std::fill(host.begin() + ps, host.begin() + 2 * ps, 0.0);
    if ((st = semx_memcpy_htod(dev, host.data(), n * sizeof(double)))) return st;
  }
  return SEMX_OK;
}

struct Header {
  std::string session, created, fields, format;
  int nr = 0, ns = 0, nz = 0, nel = 0, step = 0;
  double time = 0, dt = 0, kinvis = 0, beta = 0;
};

// The ten header lines; false at end of file or on a malformed header.
bool read_header(std::istream& in, Header& h) {
  std::string line[10];
  for (int i = 0; i < 10; ++i)
    if (!std::getline(in, line[i])) return false;
  auto value = [&](int i) { return line[i].substr(0, std::min<size_t>(25, line[i].size())); };
  {
    std::istringstream s(value(0));
    s >> h.session;
  }
  h.created = value(1);
  {
    std::istringstream s(line[2]);
    if (!(s >> h.nr >> h.ns >> h.nz >> h.nel)) return false;
  }
  std::istringstream(line[3]) >> h.step;
  std::istringstream(line[4]) >> h.time;
  std::istringstream(line[5]) >> h.dt;
  std::istringstream(line[6]) >> h.kinvis;
  std::istringstream(line[7]) >> h.beta;
  h.fields.clear();
  for (char c : line[8]) {
    if (!std::isalpha((unsigned char)c)) break;
    h.fields.push_back(c);
  }
  h.format = line[9];
  return true;
}

}  // namespace

extern "C" {

int semx_field_write(const char* path, int append, semx_mesh mesh, int nz, int nfield,
                     const char* names, double* const* fields, const char* session, int step,
                     double time, double dt, double kinvis, double beta, void* stream) {
  if (!path || !mesh || nz < 1 || nfield < 1 || !names || !fields || !session ||
      (int)std::strlen(names) != nfield)
    return SEMX_E_ARG;
  for (int i = 0; i < nfield; ++i)
    if (!fields[i] || !std::isalpha((unsigned char)names[i])) return SEMX_E_ARG;
  const semx::Mesh& M = mesh->M;
  const size_t n = (size_t)nz * M.planeSize;
  std::vector<double> host(n), out;
  double* scratch = nullptr;
  int st = (nz > 2) ? semx_malloc((void**)&scratch, n * sizeof(double)) : SEMX_OK;
  if (st) return st;
  std::ofstream f(path, append ? std::ios::binary | std::ios::app
                               : std::ios::binary | std::ios::trunc);
  if (!f) {
    if (scratch) semx_free(scratch);
    return SEMX_E_ARG;
  }
  char buf[256], date[64];
  const std::time_t tp = std::time(nullptr);
  std::strftime(date, sizeof(date), "%a %b %d %H:%M:%S %Y", std::localtime(&tp));
  std::ostringstream desc;
  desc << M.np << " " << M.np << " " << nz << " " << M.nel;
  const std::string fmt = std::string("binary IEEE ") +
                          (host_little_endian() ? "little-endian" : "big-endian");
  std::snprintf(buf, sizeof(buf), "%-25s Session\n", session); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25s Created\n", date); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25s Nr, Ns, Nz, Elements\n", desc.str().c_str()); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25d Step\n", step); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25.6g Time\n", time); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25.6g Time step\n", dt); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25.6g Kinvis\n", kinvis); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25.6g Beta\n", beta); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25s Fields written\n", names); f << buf;
  std::snprintf(buf, sizeof(buf), "%-25s Format\n", fmt.c_str()); f << buf;
  for (int i = 0; i < nfield && !st; ++i) {
    st = to_physical(M, nz, fields[i], scratch, host, out, stream);
    if (!st) f.write(reinterpret_cast<const char*>(out.data()), out.size() * sizeof(double));
  }
  if (scratch) semx_free(scratch);
  if (st) return st;
  f.flush();
  return f ? SEMX_OK : SEMX_E_ARG;
}

int semx_field_read(const char* path, int dump, semx_mesh mesh, int nz, int nfield,
                    const char* names, double* const* fields, int* step, double* time,
                    char* fields_in_file, void* stream) {
  if (!path || dump < 0 || !mesh || nz < 1 || nfield < 0 || (nfield && (!names || !fields)) ||
      (names && (int)std::strlen(names) != nfield))
    return SEMX_E_ARG;
  const semx::Mesh& M = mesh->M;
  std::ifstream f(path, std::ios::binary);
  if (!f) return SEMX_E_ARG;
  Header h;
  const size_t ntot = (size_t)M.nPlane * nz;
  for (int d = 0;; ++d) {
    if (!read_header(f, h)) return SEMX_E_ARG;
    // readField's checks (src/misc.cpp:196-205)
    if (h.nr != M.np || h.ns != M.np || h.nz != nz || h.nel != M.nel) return SEMX_E_MESH;
    if (h.format.find("binary") == std::string::npos) return SEMX_E_UNSUPPORTED;
    if (d == dump) break;
    f.seekg((std::streamoff)(h.fields.size() * ntot * sizeof(double)), std::ios::cur);
    if (!f) return SEMX_E_ARG;
  }
  const bool big = h.format.find("big") != std::string::npos;
  const bool little = h.format.find("little") != std::string::npos;
  const bool swap = (big && host_little_endian()) || (little && !host_little_endian());
  std::vector<double> phys(ntot), host((size_t)nz * M.planeSize);
  for (char c : h.fields) {
    const char* hit = names ? std::strchr(names, c) : nullptr;
    if (!hit || !*hit) {   // a field of the file that the caller did not ask for
      f.seekg((std::streamoff)(ntot * sizeof(double)), std::ios::cur);
      continue;
    }
    f.read(reinterpret_cast<char*>(phys.data()), ntot * sizeof(double));
    if (!f) return SEMX_E_ARG;
    if (swap) byte_reverse(phys.data(), phys.size());
    const int st = to_fourier(M, nz, phys.data(), fields[hit - names], host, stream);
    if (st) return st;
  }
  if (step) *step = h.step;
  if (time) *time = h.time;
  if (fields_in_file) std::snprintf(fields_in_file, 64, "%s", h.fields.c_str());
  return SEMX_OK;
}

}  // extern "C"
