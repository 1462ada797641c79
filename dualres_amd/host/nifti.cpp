// NIfTI-1 I/O and image helpers; see nifti.hpp.
#include "nifti.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace drhost {

namespace {

template <typename T>
void swap_bytes(T& v) {
  unsigned char* p = reinterpret_cast<unsigned char*>(&v);
  std::reverse(p, p + sizeof(T));
}

void swap_header(Nifti1Header& h) {
  swap_bytes(h.sizeof_hdr);
  swap_bytes(h.extents);
  swap_bytes(h.session_error);
  for (auto& d : h.dim) swap_bytes(d);
  swap_bytes(h.intent_p1);
  swap_bytes(h.intent_p2);
  swap_bytes(h.intent_p3);
  swap_bytes(h.intent_code);
  swap_bytes(h.datatype);
  swap_bytes(h.bitpix);
  swap_bytes(h.slice_start);
  for (auto& d : h.pixdim) swap_bytes(d);
  swap_bytes(h.vox_offset);
  swap_bytes(h.scl_slope);
  swap_bytes(h.scl_inter);
  swap_bytes(h.slice_end);
  swap_bytes(h.cal_max);
  swap_bytes(h.cal_min);
  swap_bytes(h.slice_duration);
  swap_bytes(h.toffset);
  swap_bytes(h.glmax);
  swap_bytes(h.glmin);
  swap_bytes(h.qform_code);
  swap_bytes(h.sform_code);
  swap_bytes(h.quatern_b);
  swap_bytes(h.quatern_c);
  swap_bytes(h.quatern_d);
  swap_bytes(h.qoffset_x);
  swap_bytes(h.qoffset_y);
  swap_bytes(h.qoffset_z);
  for (int i = 0; i < 4; ++i) {
    swap_bytes(h.srow_x[i]);
    swap_bytes(h.srow_y[i]);
    swap_bytes(h.srow_z[i]);
  }
}

// qto_xyz from the quaternion parameters (NIfTI-1 standard, method 2), or
// the pixdim scaling when qform_code is 0 (method 1).
std::array<float, 16> qform_of(const Nifti1Header& h) {
  std::array<float, 16> q{};
  if (h.qform_code <= 0) {
    q[0] = h.pixdim[1];
    q[5] = h.pixdim[2];
    q[10] = h.pixdim[3];
    q[15] = 1.0f;
    return q;
  }
  double b = h.quatern_b, c = h.quatern_c, d = h.quatern_d;
  double a = 1.0 - (b * b + c * c + d * d);
  if (a < 1.e-7) {
    a = 1.0 / std::sqrt(b * b + c * c + d * d);
    b *= a;
    c *= a;
    d *= a;
    a = 0.0;
  } else {
    a = std::sqrt(a);
  }
  const double xd = h.pixdim[1] > 0 ? h.pixdim[1] : 1.0;
  const double yd = h.pixdim[2] > 0 ? h.pixdim[2] : 1.0;
  double zd = h.pixdim[3] > 0 ? h.pixdim[3] : 1.0;
  if (h.pixdim[0] < 0) zd = -zd;
  const double R[3][3] = {{a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)},
                          {2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)},
                          {2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b}};
  const double sc[3] = {xd, yd, zd};
  const double off[3] = {h.qoffset_x, h.qoffset_y, h.qoffset_z};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) q[4 * i + j] = (float)(R[i][j] * sc[j]);
    q[4 * i + 3] = (float)off[i];
  }
  q[15] = 1.0f;
  return q;
}

struct GzFile {
  gzFile f = nullptr;
  GzFile(const std::string& p, const char* mode) : f(gzopen(p.c_str(), mode)) {}
  ~GzFile() {
    if (f) gzclose(f);
  }
};

template <typename S>
void convert(const std::vector<unsigned char>& raw, bool swap, std::vector<double>& out) {
  const size_t n = raw.size() / sizeof(S);
  out.resize(n);
  for (size_t i = 0; i < n; ++i) {
    S v;
    std::memcpy(&v, raw.data() + i * sizeof(S), sizeof(S));
    if (swap) swap_bytes(v);
    out[i] = (double)v;
  }
}

}  // namespace

bool is_nifti_file(const std::string& path) {
  GzFile g(path, "rb");
  if (!g.f) return false;
  Nifti1Header h;
  if (gzread(g.f, &h, sizeof(h)) != (int)sizeof(h)) return false;
  return (h.sizeof_hdr == 348 || __builtin_bswap32((uint32_t)h.sizeof_hdr) == 348) &&
         (std::memcmp(h.magic, "n+1", 4) == 0 || std::memcmp(h.magic, "ni1", 4) == 0);
}

Image read_nifti(const std::string& path) {
  GzFile g(path, "rb");
  if (!g.f) throw std::runtime_error("cannot open NIfTI file " + path);
  Image im;
  if (gzread(g.f, &im.hdr, sizeof(im.hdr)) != (int)sizeof(im.hdr))
    throw std::runtime_error("short NIfTI header in " + path);
  bool swap = false;
  if (im.hdr.sizeof_hdr != 348) {
    swap_header(im.hdr);
    swap = true;
    if (im.hdr.sizeof_hdr != 348) throw std::runtime_error(path + " is not a NIfTI-1 file");
  }
  if (std::memcmp(im.hdr.magic, "n+1", 4) != 0)
    throw std::runtime_error(path + ": only single-file NIfTI-1 (n+1) images are supported");
  if (im.hdr.dim[0] < 3) throw std::logic_error("NIfTI image has dim < 3");
  im.nx = im.hdr.dim[1];
  im.ny = im.hdr.dim[2];
  im.nz = im.hdr.dim[3];
  for (int k = 4; k <= im.hdr.dim[0] && k < 8; ++k)
    if (im.hdr.dim[k] > 1) throw std::runtime_error(path + ": only 3D volumes are supported");
  im.datatype = im.hdr.datatype;
  const long off = (long)im.hdr.vox_offset;
  if (off > (long)sizeof(im.hdr)) {
    std::vector<char> skip(off - sizeof(im.hdr));
    gzread(g.f, skip.data(), (unsigned)skip.size());
  }
  const int bpv = im.hdr.bitpix / 8;
  std::vector<unsigned char> raw((size_t)im.nvox() * bpv);
  size_t got = 0;
  while (got < raw.size()) {
    const unsigned chunk = (unsigned)std::min<size_t>(raw.size() - got, 1u << 30);
    const int r = gzread(g.f, raw.data() + got, chunk);
    if (r <= 0) throw std::runtime_error("short NIfTI data in " + path);
    got += (size_t)r;
  }
  switch (im.hdr.datatype) {
    case 2: convert<uint8_t>(raw, false, im.data); break;
    case 4: convert<int16_t>(raw, swap, im.data); break;
    case 8: convert<int32_t>(raw, swap, im.data); break;
    case 16: convert<float>(raw, swap, im.data); break;
    case 64: convert<double>(raw, swap, im.data); break;
    case 256: convert<int8_t>(raw, false, im.data); break;
    case 512: convert<uint16_t>(raw, swap, im.data); break;
    case 768: convert<uint32_t>(raw, swap, im.data); break;
    default: throw std::runtime_error(path + ": unsupported NIfTI datatype " + std::to_string(im.hdr.datatype));
  }
  im.qto_xyz = qform_of(im.hdr);
  return im;
}

void write_nifti(const std::string& path, const Image& like, const std::vector<double>& values, int datatype) {
  if ((long)values.size() != like.nvox()) throw std::logic_error("write_nifti: size mismatch");
  if (datatype != 16 && datatype != 64) throw std::logic_error("write_nifti: datatype must be float32 or float64");
  Nifti1Header h = like.hdr;
  h.sizeof_hdr = 348;
  h.dim[0] = 3;
  for (int k = 4; k < 8; ++k) h.dim[k] = 1;
  h.datatype = (int16_t)datatype;
  h.bitpix = datatype == 64 ? 64 : 32;
  h.vox_offset = 352.0f;
  h.scl_slope = 0.0f;
  h.scl_inter = 0.0f;
  std::memcpy(h.magic, "n+1", 4);
  const bool gz = path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
  std::remove(path.c_str());
  GzFile g(path, gz ? "wb6" : "wbT");
  if (!g.f) throw std::runtime_error("cannot write " + path);
  const char ext[4] = {0, 0, 0, 0};
  std::vector<float> f;
  if (datatype == 16) f.assign(values.begin(), values.end());
  if (gzwrite(g.f, &h, sizeof(h)) != (int)sizeof(h) || gzwrite(g.f, ext, 4) != 4)
    throw std::runtime_error("write failed: " + path);
  size_t done = 0;
  const size_t bytes = datatype == 16 ? f.size() * sizeof(float) : values.size() * sizeof(double);
  const char* p = datatype == 16 ? reinterpret_cast<const char*>(f.data()) : reinterpret_cast<const char*>(values.data());
  while (done < bytes) {
    const unsigned chunk = (unsigned)std::min<size_t>(bytes - done, 1u << 30);
    if (gzwrite(g.f, p + done, chunk) != (int)chunk) throw std::runtime_error("write failed: " + path);
    done += chunk;
  }
}

std::string nifti_basename(const std::string& path) {
  static const char* ext[] = {".nii.gz", ".hdr.gz", ".img.gz", ".nii", ".hdr", ".img"};
  for (const char* e : ext) {
    const size_t n = std::strlen(e);
    if (path.size() > n && path.compare(path.size() - n, n, e) == 0) return path.substr(0, path.size() - n);
  }
  return path;
}

bool same_orientation(const Image& a, const Image& b, float tol) {
  float m = 0.0f;
  for (int i = 0; i < 16; ++i) m = std::max(m, std::fabs(a.qto_xyz[i] - b.qto_xyz[i]));
  return m <= tol;
}

void apply_mask(Image& img, const Image& mask) {
  if (!same_orientation(img, mask)) throw std::domain_error("Image masking does not make sense: different orientations");
  if (img.nvox() != mask.nvox()) throw std::domain_error("Image masking does not make sense: different grid sizes");
  for (long i = 0; i < img.nvox(); ++i)
    if (mask.data[i] == 0.0 || std::isnan(img.data[i])) img.data[i] = 0.0;
}

void add_to(Image& a, const Image& b) {
  if (!same_orientation(a, b)) throw std::domain_error("Image math does not make sense: different orientations");
  if (a.nvox() != b.nvox()) throw std::domain_error("Image math does not make sense: different grid sizes");
  for (long i = 0; i < a.nvox(); ++i) a.data[i] += b.data[i];
}

std::vector<std::array<int, 3>> nonzero_indices(const Image& img) {
  std::vector<std::array<int, 3>> out;
  for (long i = 0; i < img.nvox(); ++i)
    if (present(img.data[i]))
      out.push_back({(int)(i % img.nx), (int)((i / img.nx) % img.ny), (int)((i / ((long)img.nx * img.ny)) % img.nz)});
  return out;
}

BBox bounding_box(const Image& img) {
  BBox b{{img.nx, img.ny, img.nz}, {-1, -1, -1}, 0};
  for (const auto& v : nonzero_indices(img)) {
    for (int a = 0; a < 3; ++a) {
      b.lo[a] = std::min(b.lo[a], v[a]);
      b.hi[a] = std::max(b.hi[a], v[a]);
    }
    ++b.nnz;
  }
  if (b.nnz == 0) throw std::domain_error("image has no nonzero voxels");
  return b;
}

std::vector<std::array<int, 3>> nonzero_indices_bounded_by_box(const Image& img, const BBox& b) {
  std::vector<std::array<int, 3>> out;
  for (auto v : nonzero_indices(img)) {
    bool ok = true;
    for (int a = 0; a < 3; ++a) ok = ok && v[a] >= b.lo[a] && v[a] <= b.hi[a];
    if (ok) {
      for (int a = 0; a < 3; ++a) v[a] -= b.lo[a];
      out.push_back(v);
    }
  }
  return out;
}

std::vector<double> nonzero_data(const Image& img) {
  std::vector<double> out;
  for (double v : img.data)
    if (present(v)) out.push_back(v);
  return out;
}

void emplace_nonzero_data(Image& img, const std::vector<double>& v) {
  size_t j = 0;
  for (long i = 0; i < img.nvox() && j < v.size(); ++i)
    if (present(img.data[i])) img.data[i] = v[j++];
}

std::array<float, 3> voxel_dimensions(const std::array<float, 16>& q) {
  std::array<float, 3> d{};
  for (int c = 0; c < 3; ++c) {
    float s = 0.0f;
    for (int r = 0; r < 4; ++r) s += q[4 * r + c] * q[4 * r + c];
    d[c] = std::sqrt(s);
  }
  return d;
}

}  // namespace drhost
