// bjenc.cpp -- TEST INFRASTRUCTURE ONLY: encodes JSON configs as BJData with the JSON library
// the reference vendors (src/nlohmann/json.hpp, built from /root/reference by `make ref`), so the
// BJData input fixtures under tests/golden/ carry that library's own encoding
// (json::to_bjdata, as src/umcx.cpp:703 calls it), and decodes BJData the way the reference reads
// its MATLAB input (json::from_bjdata, src/umcx.cpp:477), so the expected values of those fixtures
// are that library's own reading.
//   bjenc <in.json> <out.bjd> [use_size use_type]   (defaults 1 1)
//   bjenc -d <in.bjd> <out.json>                    (exit 1 and a message when it is rejected)
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "nlohmann/json.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: bjenc <in.json> <out.bjd> [use_size use_type]\n";
        return 2;
    }
    if (std::string(argv[1]) == "-d") {
        if (argc < 4) return 2;
        std::ifstream in(argv[2], std::ios::binary);
        std::vector<std::uint8_t> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        try {
            nlohmann::ordered_json j = nlohmann::ordered_json::from_bjdata(buf);
            std::ofstream(argv[3]) << j.dump();
        } catch (const std::exception& e) {
            std::cerr << e.what() << "\n";
            return 1;
        }
        return 0;
    }
    const bool use_size = argc > 3 ? std::stoi(argv[3]) != 0 : true;
    const bool use_type = argc > 4 ? std::stoi(argv[4]) != 0 : true;
    std::ifstream in(argv[1]);
    nlohmann::ordered_json j = nlohmann::ordered_json::parse(in);
    std::vector<std::uint8_t> out;
    nlohmann::ordered_json::to_bjdata(j, out, use_size, use_type);
    std::ofstream(argv[2], std::ios::binary).write((const char*)out.data(), (std::streamsize)out.size());
    return 0;
}
