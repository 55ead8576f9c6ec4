// DADA file reader (psrdada/psrdada.cpp:3-128). Errors go through msosm::check (print and exit, or
// throw in throw mode) instead of exit(1).
#include "psrdada.h"

#include <algorithm>
#include <cstring>
#include <sstream>

PSRDADA::PSRDADA(std::string filename) : infileName(std::move(filename))
{
    infile.open(infileName, std::ios::binary | std::ios::ate);
    if (!infile.is_open()) msosm::check(MSOSM_ERR_IO, ("PSRDADA: cannot open file " + infileName).c_str());
    filesize = static_cast<unsigned long>(infile.tellg());
    infile.seekg(0, std::ios::beg);
    // <name>_dedisp_folded.txt next to the input (psrdada.cpp:16)
    outfileName = infileName.substr(0, infileName.find_last_of('.')) + "_dedisp_folded.txt";
    if (filesize < (unsigned long)blockSize) msosm::check(MSOSM_ERR_IO, "PSRDADA: file shorter than the header");
    // one 4096-byte header only, as the reference reads it (psrdada.cpp:18-20)
    infile.read(headerBuffer, blockSize);
    std::string header(headerBuffer, blockSize);
    header.erase(std::remove(header.begin(), header.end(), '\0'), header.end());
    std::istringstream iss(header);
    std::string line;
    while (std::getline(iss, line)) {
        const auto kv = splitByMultipleSpaces(line);
        try {
            if (kv.first == "BW") bw = std::stof(kv.second);
            else if (kv.first == "FREQ") fc = std::stof(kv.second);
            else if (kv.first == "NBIT") nbit = std::stoi(kv.second);
            else if (kv.first == "NCHAN") nchan = std::stoi(kv.second);
            else if (kv.first == "NDIM") ndim = std::stoi(kv.second);
            else if (kv.first == "NPOL") npol = std::stoi(kv.second);
        } catch (const std::exception &) {
            msosm::check(MSOSM_ERR_IO, ("PSRDADA: bad header value in line: " + line).c_str());
        }
    }
    if (verbose) {
        std::cout << "Bandwidth: " << bw << " MHz\nCenter frequency: " << fc << " MHz\nData precision: " << nbit
                  << " bit\nChannel number: " << nchan << "\nDimension: " << ndim << "\nPolarization number: " << npol
                  << std::endl;
    }
}

PSRDADA::~PSRDADA()
{
    delete[] pol1_in;
    delete[] pol2_in;
}

void PSRDADA::initBuffer(size_t n)
{
    if (n % frameSamples != 0) msosm::check(MSOSM_ERR_ARG, "PSRDADA::initBuffer: n must be a multiple of 2048");
    if (npol != 1 && npol != 2) msosm::check(MSOSM_ERR_ARG, "PSRDADA::initBuffer: unsupported polarisation number");
    delete[] pol1_in;
    delete[] pol2_in;
    pol1_in = new uint16_pair[n];
    pol2_in = new uint16_pair[n];
    frameCount = (int)(n / frameSamples);
    readCount = static_cast<unsigned long>((filesize - blockSize) / (npol * n * sizeof(uint16_pair)));
}

void PSRDADA::readSamples()
{
    for (int i = 0; i < frameCount; i++) {
        infile.read(reinterpret_cast<char *>(pol1_in + (size_t)i * frameSamples), frameSamples * sizeof(uint16_pair));
        if (npol == 2)
            infile.read(reinterpret_cast<char *>(pol2_in + (size_t)i * frameSamples), frameSamples * sizeof(uint16_pair));
    }
    if (!infile) msosm::check(MSOSM_ERR_IO, "PSRDADA::readSamples: read past the end of the file");
}

std::pair<std::string, std::string> PSRDADA::splitByMultipleSpaces(const std::string &input)
{
    // key = text before the first space; value = the rest after the run of spaces
    const size_t pos = input.find_first_of(' ');
    if (pos == std::string::npos) return {input, ""};
    const size_t start = input.find_first_not_of(' ', pos);
    if (start == std::string::npos) return {input.substr(0, pos), ""};
    return {input.substr(0, pos), input.substr(start)};
}
