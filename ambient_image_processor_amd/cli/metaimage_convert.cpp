// metaimage_convert in.(mha|mhd) out.(mha|mhd) [--raw]
// Reads a scalar MetaImage of any supported element type and byte order
// (compressed or not) and writes it as MET_FLOAT (zlib-compressed unless
// --raw): the conversion the deshade CLI applies to its input.
#include "metaimage.hpp"

#include <cstring>
#include <iostream>

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: metaimage_convert in.mha out.mha [--raw]" << std::endl;
        return 1;
    }
    try {
        const metaimage::Volume v = metaimage::read(argv[1]);
        metaimage::write(argv[2], v, !(argc > 3 && std::strcmp(argv[3], "--raw") == 0));
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
