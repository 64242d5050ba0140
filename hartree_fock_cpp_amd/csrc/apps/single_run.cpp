// single_run -- one RHF or UHF calculation from a JSON input
// (the reference's apps/single_run.cpp: same usage, output folder and errors).
#include <exception>
#include <iostream>

#include "../host/run.hpp"

int main(int argc, char** argv) {
    try {
        if (argc < 2) {
            std::cerr << "Usage: " << argv[0] << " input.json\n";
            return 1;
        }
        hfxh::RunOptions opt = hfxh::read_scf_options(argv[1]);
        hfxh::run_scf_from_options(opt);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << "\n";
        return 1;
    }
}
