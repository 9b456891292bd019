#include "csvIO.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace {

bool ends_csv(const std::string& s) { return s.find(".csv") != std::string::npos; }

std::vector<std::string> split(const std::string& line) {
  std::vector<std::string> out;
  std::string item;
  std::istringstream is(line);
  while (std::getline(is, item, ',')) out.push_back(item);
  return out;
}

// std::stod semantics: leading blanks skipped, trailing text ignored, no number = error.
double to_double(const std::string& s) {
  const char* b = s.c_str();
  char* e = nullptr;
  errno = 0;
  double v = std::strtod(b, &e);
  if (e == b) throw std::invalid_argument("stod");
  if (errno == ERANGE && (v == HUGE_VAL || v == -HUGE_VAL)) throw std::out_of_range("stod");
  return v;
}

}  // namespace

csvReader::csvReader(std::shared_ptr<Context> ctx, std::string filename)
    : ctx_(std::move(ctx)), filename_(std::move(filename)) {
  if (!ends_csv(filename_)) {
    std::string msg = "Input file should end with '.csv', actual: " + filename_;
    ctx_->getLogger()->error(msg);
    throw std::invalid_argument(msg);
  }
}

csvReader::HeaderInfo csvReader::inspectHeader(const std::string& header) {
  HeaderInfo h;
  for (const std::string& c : split(header)) {
    if (c.find("CO") != std::string::npos) ++h.nsp;
    if (c.find("ENTHA") != std::string::npos) h.hasEnthalpy = true;
    if (c.find("TEMPE") != std::string::npos) h.hasTemperature = true;
    if (c.find("Points") != std::string::npos) h.hasCoords = true;
  }
  return h;
}

void csvReader::read() {
  std::ifstream f(filename_);
  if (!f.is_open()) {
    std::string msg = "Failing opening input file: " + filename_;
    ctx_->getLogger()->error(msg);
    throw std::runtime_error(msg);
  }
  ctx_->getLogger()->info("Reading file: \"" + filename_ + "\"");
  std::string line;
  std::getline(f, line);
  const HeaderInfo h = inspectHeader(line);
  Mesh& mesh = *ctx_->getMesh();
  while (std::getline(f, line)) {
    std::vector<std::string> it = split(line);
    size_t k = 0;
    auto next = [&]() -> double {
      if (k >= it.size()) throw std::invalid_argument("stod");
      return to_double(it[k++]);
    };
    std::vector<double> Y(h.nsp);
    for (double& y : Y) y = next();
    double hval = h.hasEnthalpy ? next() : 0.0;
    if (!h.hasTemperature) throw std::runtime_error("Input file does not have temperature!!");
    double T = next();
    PointBuilder pb(h.nsp);
    pb.addSpecies(std::move(Y)).addTemperature(T);
    if (h.hasEnthalpy) pb.addEnthalpy(hval);
    if (h.hasCoords) {
      Coords c;
      c.x = next();
      c.y = next();
      c.z = next();
      pb.addCoordinates(c);
    }
    mesh.addPoint(pb);
  }
  ctx_->getLogger()->info("Read completed");
}

csvWriter::csvWriter(std::shared_ptr<Context> ctx, std::string filename)
    : ctx_(std::move(ctx)), filename_(std::move(filename)) {
  if (!ends_csv(filename_)) {
    std::string msg = "Output file should end with '.csv', actual: " + filename_;
    ctx_->getLogger()->error(msg);
    throw std::invalid_argument(msg);
  }
}

std::string csvWriter::path() const {
  const std::string& d = ctx_->outFolder;
  if (d.empty() || d == ".") return filename_;
  return d.back() == '/' ? d + filename_ : d + "/" + filename_;
}

void csvWriter::write() {
  const std::string p = path();
  std::FILE* f = std::fopen(p.c_str(), "w");
  if (!f) {
    std::string msg = "Failing opening output file: " + p;
    ctx_->getLogger()->error(msg);
    throw std::runtime_error(msg);
  }
  Mesh& mesh = *ctx_->getMesh();
  const int nsp = mesh.numSpecies();
  // header: the first formatted item is padded to width 3 with '0'
  std::string head;
  for (int i = 0; i < nsp; ++i) head += "CO" + std::to_string(i) + ",";
  if (mesh.hasEnthalpy()) head += "ENTHA,";
  head += "TEMPE";
  if (head.size() >= 2 && head.compare(0, 2, "CO") == 0) head = "0" + head;
  if (mesh.hasCoordinates()) head += ",Points:0,Points:1,Points:2";
  std::fputs(head.c_str(), f);
  std::fputc('\n', f);
  const int n = mesh.totalSize();
  for (int c = 0; c < n; ++c) {
    const std::vector<double>& Y = mesh.getSpeciesVector(c);
    for (int i = 0; i < nsp; ++i) std::fprintf(f, "%e,", Y[i]);
    if (mesh.hasEnthalpy()) std::fprintf(f, "%e,", mesh.getEnthalpyVector()[c]);
    std::fprintf(f, "%g", mesh.getTemperatureVector()[c]);
    if (mesh.hasCoordinates()) {
      const Coords& x = mesh.getCoordinatesVector()[c];
      std::fprintf(f, ",%g,%g,%g", x.x, x.y, x.z);
    }
    std::fputc('\n', f);
  }
  if (std::fclose(f) != 0) throw std::runtime_error("Failing writing output file: " + p);
  ctx_->getLogger()->info("Write completed");
}
