// Mesh file reader and writer in the reference's CSV format
// (src/InputOutput/Reader/csvReader.cpp:14-160, src/InputOutput/Writer/csvWriter.cpp:15-111).
//
// Input: a header line, then one line per cell: nsp species mass fractions, optional
// enthalpy, temperature, optional x,y,z. Columns are recognised by substring: "CO" (a
// species), "ENTHA", "TEMPE", "Points".
// Output: header "0CO0,CO1,...,CO{nsp-1}[,ENTHA],TEMPE[,Points:0,Points:1,Points:2]" (the
// leading 0 is the reference writer's setw(3)/setfill('0') applied to "CO"); species and
// enthalpy in %e with 6 digits, temperature and coordinates in %g with 6 significant digits.
// The writer's text is lossy (6 digits); compare results in memory, not through it.
#pragma once

#include <memory>
#include <string>

#include "Context.hpp"

class csvReader {
 public:
  // Throws std::invalid_argument("Input file should end with '.csv', actual: <name>").
  csvReader(std::shared_ptr<Context> ctx, std::string filename);
  // Appends the file's cells to the context's mesh. Throws std::runtime_error
  // ("Failing opening input file: <name>", "Input file does not have temperature!!").
  void read();

  struct HeaderInfo {
    int nsp = 0;
    bool hasTemperature = false, hasEnthalpy = false, hasCoords = false;
  };
  static HeaderInfo inspectHeader(const std::string& header);

 private:
  std::shared_ptr<Context> ctx_;
  std::string filename_;
};

class csvWriter {
 public:
  // Throws std::invalid_argument("Output file should end with '.csv', actual: <name>").
  csvWriter(std::shared_ptr<Context> ctx, std::string filename);
  // Writes the mesh to <ctx->outFolder>/<filename>. Throws std::runtime_error
  // ("Failing opening output file: <path>").
  void write();
  std::string path() const;

 private:
  std::shared_ptr<Context> ctx_;
  std::string filename_;
};
