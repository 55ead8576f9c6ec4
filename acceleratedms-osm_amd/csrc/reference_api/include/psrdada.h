// PSRDADA: reader of DADA baseband files (psrdada/include/psrdada.h, psrdada/psrdada.cpp): a 4096-byte
// ASCII header of "KEY value" lines, then frames of 2048 uint16-pair samples, one frame per
// polarisation in turn.
#pragma once

#include <fstream>
#include <string>
#include <utility>

#include "globals.h"

constexpr int blockSize = 4096;
constexpr int frameSamples = 2048;

class PSRDADA {
public:
    explicit PSRDADA(std::string filename);
    ~PSRDADA();
    // buffers of n samples per polarisation (n a multiple of frameSamples)
    void initBuffer(size_t n);
    // the next n samples of each polarisation
    void readSamples();

private:
    std::pair<std::string, std::string> splitByMultipleSpaces(const std::string &input);

public:
    char headerBuffer[blockSize];
    std::ifstream infile;
    std::string infileName;
    unsigned long filesize = 0;
    unsigned long readCount = 0;
    std::string outfileName;
    float bw = 0;   // MHz
    float fc = 0;   // MHz (band centre)
    int nbit = 0;
    int nchan = 0;
    int ndim = 0;
    int npol = 0;
    int hdr_size = blockSize;
    uint16_pair *pol1_in = nullptr;
    uint16_pair *pol2_in = nullptr;
    int frameCount = 0;
};
