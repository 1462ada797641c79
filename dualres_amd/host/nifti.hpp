// Minimal NIfTI-1 reader/writer (.nii and .nii.gz via zlib) and the image
// helpers the sampler path needs.  Replaces the vendored nifticlib calls and
// include/dualres/impl/nifti_manipulation.inl for the dualgpm / estimate_rbf
// tools.  Volumes are held as float64 in voxel order (x fastest); 0 and NaN
// mark missing voxels, as in the reference.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace drhost {

#pragma pack(push, 1)
struct Nifti1Header {  // NIfTI-1 header layout (348 bytes)
  int32_t sizeof_hdr;
  char data_type[10];
  char db_name[18];
  int32_t extents;
  int16_t session_error;
  char regular;
  char dim_info;
  int16_t dim[8];
  float intent_p1, intent_p2, intent_p3;
  int16_t intent_code, datatype, bitpix, slice_start;
  float pixdim[8];
  float vox_offset, scl_slope, scl_inter;
  int16_t slice_end;
  char slice_code, xyzt_units;
  float cal_max, cal_min, slice_duration, toffset;
  int32_t glmax, glmin;
  char descrip[80];
  char aux_file[24];
  int16_t qform_code, sform_code;
  float quatern_b, quatern_c, quatern_d, qoffset_x, qoffset_y, qoffset_z;
  float srow_x[4], srow_y[4], srow_z[4];
  char intent_name[16];
  char magic[4];
};
#pragma pack(pop)
static_assert(sizeof(Nifti1Header) == 348, "NIfTI-1 header must be 348 bytes");

struct Image {
  Nifti1Header hdr{};
  int nx = 0, ny = 0, nz = 0;
  std::vector<double> data;         // nx*ny*nz, x fastest
  std::array<float, 16> qto_xyz{};  // row-major 4x4 (nifti_manipulation.inl:543-551)
  int datatype = 16;
  long nvox() const { return (long)nx * ny * nz; }
};

bool is_nifti_file(const std::string& path);
Image read_nifti(const std::string& path);
// Writes `values` (nvox, voxel order) with the geometry of `like`, as float32
// (datatype 16) or float64 (datatype 64).
void write_nifti(const std::string& path, const Image& like, const std::vector<double>& values, int datatype = 16);
// nifti_makebasename: the path without a .nii/.hdr/.img extension (and .gz)
std::string nifti_basename(const std::string& path);

inline bool present(double v) { return !(v != v || v == 0.0); }

// nifti_manipulation.inl equivalents
bool same_orientation(const Image& a, const Image& b, float tol = 1e-4f);
void apply_mask(Image& img, const Image& mask);          // :113-125
void add_to(Image& a, const Image& b);                   // :55-66
std::vector<std::array<int, 3>> nonzero_indices(const Image& img);  // :334-360
struct BBox { std::array<int, 3> lo, hi; long nnz; };
BBox bounding_box(const Image& img);                     // :210-219
std::vector<std::array<int, 3>> nonzero_indices_bounded_by_box(const Image& img, const BBox& b);  // :377-398
std::vector<double> nonzero_data(const Image& img);      // :297-314
void emplace_nonzero_data(Image& img, const std::vector<double>& v);  // :185-201
std::array<float, 3> voxel_dimensions(const std::array<float, 16>& q);  // :600-605

}  // namespace drhost
