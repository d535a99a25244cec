// Command-line renderer with the interface of lajolla's (src/main.cpp):
//   lajolla_amd [-d device] [-o output.pfm] scene.xml ...
// Each scene is parsed, rendered on the device and written as a PFM image.
#include "lajolla_amd.hpp"

#include <chrono>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char *argv[]) {
    if (argc <= 1) {
        std::cout << "[Usage] " << argv[0] << " [-d device] [-o output_file_name.pfm] filename.xml" << std::endl;
        return 0;
    }
    int device = 0;
    std::string output;
    std::vector<std::string> files;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if ((a == "-d" || a == "-t") && i + 1 < argc) {
            int v = std::stoi(argv[++i]);
            if (a == "-d") device = v;  // -t (host threads) is accepted and unused: the device renders
        } else if (a == "-o" && i + 1 < argc) {
            output = argv[++i];
        } else {
            files.push_back(a);
        }
    }
    try {
        for (const std::string &f : files) {
            auto t0 = std::chrono::steady_clock::now();
            std::cout << "Parsing and constructing scene " << f << "." << std::endl;
            auto scene = lajolla_amd::parse_scene(f, device);
            auto t1 = std::chrono::steady_clock::now();
            std::cout << "Done. Took " << std::chrono::duration<double>(t1 - t0).count() << " seconds." << std::endl;
            std::cout << "Rendering..." << std::endl;
            lajolla_amd::Image3 img = lajolla_amd::render(*scene);
            auto t2 = std::chrono::steady_clock::now();
            std::cout << "Done. Took " << std::chrono::duration<double>(t2 - t1).count() << " seconds." << std::endl;
            std::string out = output.empty() ? scene->output_filename : output;
            lajolla_amd::imwrite(out, img);
            std::cout << "Image written to " << out << std::endl;
        }
    } catch (const lajolla_amd::Error &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
