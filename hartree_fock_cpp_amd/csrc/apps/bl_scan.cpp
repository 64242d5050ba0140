// bl_scan -- RHF bond-length scan from a JSON input
// (the reference's apps/bl_scan.cpp: same usage, output folder and errors).
#include <exception>
#include <iostream>

#include "../host/run.hpp"

int main(int argc, char** argv) {
    try {
        if (argc != 2) {
            std::cerr << "Usage: " << argv[0] << " <input.json>\n";
            return 1;
        }
        std::cout << "Reading input file...\n";
        hfxh::BlScanOptions opt = hfxh::read_blscan_options(argv[1]);
        std::cout << "Starting bond-length scan...\n";
        hfxh::run_bl_scan(opt);
        std::cout << "Bond-length scan completed successfully.\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
