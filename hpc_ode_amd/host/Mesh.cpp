#include "Mesh.hpp"

#include <stdexcept>
#include <string>

PointBuilder Point::create(int nsp) { return PointBuilder(nsp); }

PointBuilder& PointBuilder::addTemperature(double T) {
  p_.T_ = T;
  p_.has_T_ = true;
  return *this;
}

PointBuilder& PointBuilder::addEnthalpy(double h) {
  p_.h_ = h;
  p_.has_h_ = true;
  return *this;
}

PointBuilder& PointBuilder::addSpecies(std::vector<double> Y) {
  if (static_cast<int>(Y.size()) != p_.nsp_)
    throw std::invalid_argument("species vector of length " + std::to_string(Y.size()) +
                                " for a point of " + std::to_string(p_.nsp_) + " species");
  p_.Y_ = std::move(Y);
  p_.has_Y_ = true;
  return *this;
}

PointBuilder& PointBuilder::addCoordinates(Coords c) {
  p_.xyz_ = c;
  p_.has_xyz_ = true;
  return *this;
}

std::vector<double>& Mesh::getEnthalpyVector() {
  if (!has_h_) throw std::runtime_error("Mesh does not have enthalpy");
  return enthalpy_;
}

std::vector<Coords>& Mesh::getCoordinatesVector() {
  if (!has_xyz_) throw std::runtime_error("Mesh does not have coordinates");
  return coords_;
}

void Mesh::addPoint(const Point& p) {
  if (temperature_.empty() && p.isReady()) {
    nsp_ = p.numSpecies();
    has_h_ = p.hasEnthalpy();
    has_xyz_ = p.hasCoordinates();
  }
  const bool compatible = p.isReady() && p.numSpecies() == nsp_ && p.hasEnthalpy() == has_h_ &&
                          p.hasCoordinates() == has_xyz_;
  if (!compatible) throw std::runtime_error("Trying to add an incompatible Point to Mesh");
  temperature_.push_back(p.getTemperature());
  species_.push_back(p.getSpecies());
  if (has_h_) enthalpy_.push_back(p.getEnthalpy());
  if (has_xyz_) coords_.push_back(p.getCoordinates());
}

void Mesh::clear() {
  temperature_.clear();
  enthalpy_.clear();
  species_.clear();
  coords_.clear();
  nsp_ = 0;
  has_h_ = has_xyz_ = false;
}
