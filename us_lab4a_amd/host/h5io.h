// HDF5 files in the reference's layouts (HDF5 C library).
//
// Acquisitions: one file per transmit, <dir>/signals-base%04d-tx%04d.h5
// (FileUtil.cpp:50-65 txElemSignalsPath), dataset "signal" [rx][samples]
// (SavedSTAAcquisition.h:78-90).  Images: <dir>/image_value.h5 "value",
// image_factor.h5 "factor", image_cf.h5 "cf", image_x.h5 "x", image_z.h5
// "z", each [n1][n2] (Project.h:106-113).  Datasets are written as
// IEEE_F64LE, chunked and deflated at level 1 like HDF5Util::save
// (HDF5Util.h:260-290); reads convert any stored float type.
#ifndef MISTA_H5IO_H
#define MISTA_H5IO_H

#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace mista_h5 {

inline std::string tx_signals_path(const std::string& dir, unsigned base, unsigned tx)
{
	char name[64];
	std::snprintf(name, sizeof(name), "/signals-base%04u-tx%04u.h5", base, tx);
	return dir + name;
}

struct Closer {
	hid_t id;
	herr_t (*close)(hid_t);
	~Closer() { if (id >= 0) close(id); }
};

inline std::vector<float> read2d(const std::string& path, const std::string& name, size_t& n1, size_t& n2)
{
	const hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0) throw std::runtime_error("cannot open " + path);
	Closer cf{f, H5Fclose};
	const hid_t d = H5Dopen2(f, name.c_str(), H5P_DEFAULT);
	if (d < 0) throw std::runtime_error("no dataset " + name + " in " + path);
	Closer cd{d, H5Dclose};
	const hid_t sp = H5Dget_space(d);
	Closer cs{sp, H5Sclose};
	if (H5Sget_simple_extent_ndims(sp) != 2) throw std::runtime_error("dataset " + name + " in " + path + " is not 2-D");
	hsize_t dims[2];
	H5Sget_simple_extent_dims(sp, dims, nullptr);
	n1 = dims[0];
	n2 = dims[1];
	std::vector<float> v(n1 * n2);
	if (H5Dread(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
		throw std::runtime_error("cannot read " + name + " from " + path);
	return v;
}

inline void write2d(const std::string& path, const std::string& name, const float* data, size_t n1, size_t n2)
{
	const hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	if (f < 0) throw std::runtime_error("cannot create " + path);
	Closer cf{f, H5Fclose};
	const hsize_t dims[2] = {n1, n2};
	const hid_t sp = H5Screate_simple(2, dims, nullptr);
	Closer cs{sp, H5Sclose};
	const hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
	Closer cp{dcpl, H5Pclose};
	if (H5Zfilter_avail(H5Z_FILTER_DEFLATE) > 0 && n1 * n2 > 0) {
		const hsize_t chunk[2] = {std::min<hsize_t>(n1, 64), std::min<hsize_t>(n2, 4096)};
		H5Pset_chunk(dcpl, 2, chunk);
		H5Pset_deflate(dcpl, 1);
	}
	const hid_t d = H5Dcreate2(f, name.c_str(), H5T_IEEE_F64LE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
	if (d < 0) throw std::runtime_error("cannot create dataset " + name + " in " + path);
	Closer cd{d, H5Dclose};
	if (H5Dwrite(d, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
		throw std::runtime_error("cannot write " + name + " to " + path);
}

}  // namespace mista_h5

#endif
