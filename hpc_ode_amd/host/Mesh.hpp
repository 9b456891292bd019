// Cell store of the integration: the same interface as the reference's Mesh, Point and
// PointBuilder (include/ODEIntegrator/Mesh/Mesh.hpp:8-44, Point.hpp:5-48,
// PointBuilder.hpp:8-27), so an integrator written against one compiles against the other.
//
// Layout (as in the reference): temperature[N], enthalpy[N] (optional), species[N][nsp]
// with all nsp mass fractions of a cell contiguous, coordinates[N] (optional). The
// integrator mutates temperature and species in place.
#pragma once

#include <vector>

struct Coords {
  double x, y, z;
};

class PointBuilder;

// One cell; built through PointBuilder (Point.hpp:13-14 of the reference).
class Point {
 public:
  friend class PointBuilder;
  static PointBuilder create(int nsp);

  bool hasTemperature() const { return has_T_; }
  bool hasEnthalpy() const { return has_h_; }
  bool hasSpecies() const { return has_Y_; }
  bool hasCoordinates() const { return has_xyz_; }
  // A point can enter a Mesh once it carries temperature and species (Point.cpp of the reference).
  bool isReady() const { return has_T_ && has_Y_; }

  double getTemperature() const { return T_; }
  double getEnthalpy() const { return h_; }
  std::vector<double> getSpecies() const { return Y_; }
  Coords getCoordinates() const { return xyz_; }
  int numSpecies() const { return nsp_; }

 private:
  explicit Point(int nsp) : nsp_(nsp) {}
  double T_ = 0.0, h_ = 0.0;
  std::vector<double> Y_;
  Coords xyz_{0.0, 0.0, 0.0};
  int nsp_;
  bool has_T_ = false, has_h_ = false, has_Y_ = false, has_xyz_ = false;
};

class PointBuilder {
 public:
  explicit PointBuilder(int nsp) : p_(Point(nsp)) {}
  operator Point() const { return p_; }
  // Each field can be set once (std::logic_error otherwise). addTemperature rejects T < 0;
  // addSpecies rejects a vector whose length differs from nsp, an entry above 1 or a sum
  // outside (0.99, 1.01) (std::runtime_error, the reference's PointBuilder.cpp:8-70).
  PointBuilder& addTemperature(double T);
  PointBuilder& addEnthalpy(double h);
  PointBuilder& addSpecies(std::vector<double> Y);
  PointBuilder& addCoordinates(Coords c);

 private:
  Point p_;
};

class Mesh {
 public:
  Mesh() = default;
  Mesh(const Mesh&) = delete;
  Mesh& operator=(const Mesh&) = delete;

  std::vector<double>& getTemperatureVector() { return temperature_; }
  std::vector<double>& getEnthalpyVector();                 // throws when the mesh has no enthalpy
  std::vector<double>& getSpeciesVector(int index) { return species_.at(index); }
  std::vector<std::vector<double>>& getSpeciesMatrix() { return species_; }
  std::vector<Coords>& getCoordinatesVector();              // throws when the mesh has no coordinates

  double* getTemperaturePointer() { return temperature_.data(); }
  double* getEnthalpyPointer() { return getEnthalpyVector().data(); }
  double* getSpeciesPointer(int index) { return species_.at(index).data(); }

  // Appends a cell; the first ready point fixes nsp and which optional fields the mesh
  // carries. Throws std::runtime_error("Trying to add an incompatible Point to Mesh")
  // otherwise (Mesh.cpp:51-71 of the reference).
  void addPoint(const Point& p);
  int totalSize() const { return static_cast<int>(temperature_.size()); }
  int numSpecies() const { return nsp_; }
  bool hasEnthalpy() const { return has_h_; }
  bool hasCoordinates() const { return has_xyz_; }
  void clear();

 private:
  std::vector<double> temperature_, enthalpy_;
  std::vector<std::vector<double>> species_;
  std::vector<Coords> coords_;
  int nsp_ = 0;
  bool has_h_ = false, has_xyz_ = false;
};
