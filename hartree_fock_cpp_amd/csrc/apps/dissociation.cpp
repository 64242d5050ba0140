// dissociation -- RHF molecule + UHF atoms dissociation energy from a JSON
// input (the reference's apps/dissociation.cpp: same usage, output and errors).
#include <exception>
#include <iostream>

#include "../host/run.hpp"

int main(int argc, char** argv) {
    try {
        if (argc != 2) {
            std::cerr << "Usage: " << argv[0] << " <input.json>\n";
            return 1;
        }
        hfxh::RunOptions opt = hfxh::read_scf_options(argv[1]);
        hfxh::run_dissociation(opt);
        std::cout << "Dissociation calculation completed.\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
