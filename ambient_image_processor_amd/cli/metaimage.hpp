// Minimal MetaImage (.mha / .mhd + raw) reader and writer for 2-D/3-D scalar
// volumes, the format the reference reads and writes through ITK
// (itk/ITKImage.cpp:116-147 read, :206-232 write with compression on).
//
// Read: ElementType MET_{UCHAR,CHAR,USHORT,SHORT,UINT,INT,ULONG,LONG,
// ULONG_LONG,LONG_LONG,FLOAT,DOUBLE}, either byte order, uncompressed or zlib
// compressed, ElementDataFile LOCAL or a file next to the header; converted to
// float (ITKImage::PixelType).  Write: MET_FLOAT, little endian, LOCAL for .mha
// or a .raw/.zraw file for .mhd, zlib-compressed by default (as the
// reference's writer->UseCompressionOn()), keeping spacing and origin.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace metaimage {

struct Volume {
    uint32_t w = 0, h = 0, d = 1;
    double spacing[3] = {1.0, 1.0, 1.0};
    double origin[3] = {0.0, 0.0, 0.0};
    std::vector<float> data;   // x fastest
};

inline std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

inline std::string lower(std::string s) {
    for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

inline bool ends_with(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && lower(s.substr(s.size() - suf.size())) == suf;
}

inline std::string dir_of(const std::string& path) {
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

inline size_t type_size(const std::string& t) {
    if (t == "MET_UCHAR" || t == "MET_CHAR") return 1;
    if (t == "MET_USHORT" || t == "MET_SHORT") return 2;
    if (t == "MET_UINT" || t == "MET_INT" || t == "MET_FLOAT") return 4;
    if (t == "MET_ULONG" || t == "MET_LONG" || t == "MET_ULONG_LONG" || t == "MET_LONG_LONG" ||
        t == "MET_DOUBLE")
        return 8;
    throw std::runtime_error("unsupported ElementType " + t);
}

template <typename T>
inline T load_le(const unsigned char* p, bool msb) {
    unsigned char b[sizeof(T)];
    for (size_t i = 0; i < sizeof(T); ++i) b[i] = msb ? p[sizeof(T) - 1 - i] : p[i];
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

inline float convert(const unsigned char* p, const std::string& t, bool msb) {
    if (t == "MET_UCHAR") return float(*p);
    if (t == "MET_CHAR") return float(int8_t(*p));
    if (t == "MET_USHORT") return float(load_le<uint16_t>(p, msb));
    if (t == "MET_SHORT") return float(load_le<int16_t>(p, msb));
    if (t == "MET_UINT") return float(load_le<uint32_t>(p, msb));
    if (t == "MET_INT") return float(load_le<int32_t>(p, msb));
    if (t == "MET_FLOAT") return load_le<float>(p, msb);
    if (t == "MET_DOUBLE") return float(load_le<double>(p, msb));
    if (t == "MET_ULONG" || t == "MET_ULONG_LONG") return float(load_le<uint64_t>(p, msb));
    return float(load_le<int64_t>(p, msb));
}

// zlib counts bytes per call in 32 bits (uInt): payloads of 4 GiB and more
// are fed and drained in pieces of at most `chunk` bytes.
inline std::vector<unsigned char> inflate_all(const std::vector<unsigned char>& in, size_t out_size,
                                              size_t chunk = size_t(UINT_MAX)) {
    std::vector<unsigned char> out(out_size);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) throw std::runtime_error("zlib inflateInit failed");
    unsigned char spare = 0;   // next_out target once the output is full
    zs.next_out = out_size ? out.data() : &spare;
    size_t in_pos = 0, out_pos = 0;
    int r = Z_OK;
    while (r != Z_STREAM_END) {
        if (zs.avail_in == 0 && in_pos < in.size()) {
            const size_t n = std::min(chunk, in.size() - in_pos);
            zs.next_in = const_cast<unsigned char*>(in.data()) + in_pos;
            zs.avail_in = uInt(n);
            in_pos += n;
        }
        if (zs.avail_out == 0 && out_pos < out_size) {
            const size_t n = std::min(chunk, out_size - out_pos);
            zs.next_out = out.data() + out_pos;
            zs.avail_out = uInt(n);
            out_pos += n;
        }
        // with the output full inflate still reads the stream's trailer; it
        // returns Z_BUF_ERROR when no progress is possible (truncated input,
        // or more output than the header declares)
        r = inflate(&zs, Z_NO_FLUSH);
        if (r != Z_OK && r != Z_STREAM_END) break;
    }
    const bool ok = r == Z_STREAM_END && size_t(zs.total_out) == out_size;
    inflateEnd(&zs);
    if (!ok) throw std::runtime_error("compressed element data is corrupt or truncated");
    return out;
}

inline Volume read(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line, type, datafile;
    int ndims = 0, channels = 1;
    bool msb = false, compressed = false;
    size_t compressed_size = 0;
    std::vector<uint32_t> dims;
    Volume v;
    while (std::getline(f, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        std::istringstream vs(val);
        if (key == "NDims") vs >> ndims;
        else if (key == "DimSize") { uint32_t x; while (vs >> x) dims.push_back(x); }
        else if (key == "ElementType") type = val;
        else if (key == "BinaryDataByteOrderMSB" || key == "ElementByteOrderMSB") msb = lower(val) == "true";
        else if (key == "CompressedData") compressed = lower(val) == "true";
        else if (key == "CompressedDataSize") vs >> compressed_size;
        else if (key == "ElementNumberOfChannels") vs >> channels;
        else if (key == "ElementSpacing" || key == "ElementSize") {
            for (int i = 0; i < 3 && (vs >> v.spacing[i]); ++i) {}
        } else if (key == "Offset" || key == "Origin" || key == "Position") {
            for (int i = 0; i < 3 && (vs >> v.origin[i]); ++i) {}
        } else if (key == "ElementDataFile") { datafile = val; break; }
    }
    if (datafile.empty()) throw std::runtime_error("no ElementDataFile in " + path);
    if (ndims < 2 || ndims > 3 || dims.size() != size_t(ndims))
        throw std::runtime_error("only 2-D and 3-D images are supported");
    if (channels != 1) throw std::runtime_error("only scalar images are supported");
    v.w = dims[0]; v.h = dims[1]; v.d = ndims == 3 ? dims[2] : 1;
    const size_t n = size_t(v.w) * v.h * v.d, es = type_size(type), raw_size = n * es;
    std::vector<unsigned char> bytes;
    auto slurp = [&](std::istream& s) {
        std::vector<unsigned char> b((std::istreambuf_iterator<char>(s)), std::istreambuf_iterator<char>());
        return b;
    };
    if (datafile == "LOCAL") {
        bytes = slurp(f);
    } else {
        std::ifstream df(dir_of(path) + datafile, std::ios::binary);
        if (!df) throw std::runtime_error("cannot open element data file " + datafile);
        bytes = slurp(df);
    }
    if (compressed) {
        if (compressed_size && compressed_size < bytes.size()) bytes.resize(compressed_size);
        bytes = inflate_all(bytes, raw_size);
    }
    if (bytes.size() < raw_size) throw std::runtime_error("element data is truncated");
    v.data.resize(n);
    for (size_t i = 0; i < n; ++i) v.data[i] = convert(bytes.data() + i * es, type, msb);
    return v;
}

inline void write(const std::string& path, const Volume& v, bool compress = true) {
    const bool mhd = ends_with(path, ".mhd");
    if (!mhd && !ends_with(path, ".mha"))
        throw std::runtime_error("output must be a MetaImage file (.mha or .mhd): " + path);
    const size_t n = size_t(v.w) * v.h * v.d;
    std::vector<unsigned char> payload(n * 4);
    std::memcpy(payload.data(), v.data.data(), n * 4);   // little-endian hosts
    if (compress) {
        uLongf cap = compressBound(uLong(payload.size()));
        std::vector<unsigned char> z(cap);
        if (compress2(z.data(), &cap, payload.data(), uLong(payload.size()), Z_DEFAULT_COMPRESSION) != Z_OK)
            throw std::runtime_error("zlib compression failed");
        z.resize(cap);
        payload.swap(z);
    }
    const int nd = v.d > 1 ? 3 : 2;
    std::ostringstream h;
    h << "ObjectType = Image\nNDims = " << nd << "\nBinaryData = True\n"
      << "BinaryDataByteOrderMSB = False\nCompressedData = " << (compress ? "True" : "False") << "\n";
    if (compress) h << "CompressedDataSize = " << payload.size() << "\n";
    h.precision(17);
    h << "Offset =";
    for (int i = 0; i < nd; ++i) h << " " << v.origin[i];
    h << "\nElementSpacing =";
    for (int i = 0; i < nd; ++i) h << " " << v.spacing[i];
    h << "\nDimSize = " << v.w << " " << v.h;
    if (nd == 3) h << " " << v.d;
    h << "\nElementType = MET_FLOAT\n";
    std::string data_name;
    if (mhd) {
        std::string base = path.substr(0, path.size() - 4);
        const size_t slash = base.find_last_of('/');
        data_name = (slash == std::string::npos ? base : base.substr(slash + 1)) +
                    (compress ? ".zraw" : ".raw");
        h << "ElementDataFile = " << data_name << "\n";
    } else {
        h << "ElementDataFile = LOCAL\n";
    }
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot write " + path);
    const std::string hs = h.str();
    out.write(hs.data(), std::streamsize(hs.size()));
    if (mhd) {
        std::ofstream raw(dir_of(path) + data_name, std::ios::binary);
        if (!raw) throw std::runtime_error("cannot write " + data_name);
        raw.write(reinterpret_cast<const char*>(payload.data()), std::streamsize(payload.size()));
    } else {
        out.write(reinterpret_cast<const char*>(payload.data()), std::streamsize(payload.size()));
    }
}

}  // namespace metaimage
