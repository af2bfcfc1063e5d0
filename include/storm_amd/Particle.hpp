// Particle record, status codes and step result of the drop-in interface.
//
// Field for field the reference's packet record (particle/Particle.hpp:67-85,
// non-MPI build), its status enum (particle/ParticleStatus.hpp:8-15) and
// step result (particle/StepResult.hpp:9-14), so that user code written
// against the reference compiles unchanged.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <ostream>
#include <string>

namespace STORM {

using dt_t = double;

enum ParticleStatus { NO_CELL_MOVE, CELL_MOVE, DONE, REMOVE, REFLECT };

inline std::string ParticleStatusToString(int status) {
  switch (status) {
    case NO_CELL_MOVE: return "NO_CELL_MOVE";
    case CELL_MOVE: return "CELL_MOVE";
    case DONE: return "DONE";
    case REMOVE: return "REMOVE";
    case REFLECT: return "REFLECT";
    default: return "UNKNOWN(" + std::to_string(status) + ")";
  }
}

template <typename T, typename Grid>
struct StepResult {
  ParticleStatus change = ParticleStatus::NO_CELL_MOVE;
  size_t nextCellIndex = std::numeric_limits<size_t>::max();
};

template <typename T, typename Grid>
struct Particle {
  size_t id = std::numeric_limits<size_t>::max();
  size_t cellID = std::numeric_limits<size_t>::max();
  size_t sourceCellID = std::numeric_limits<size_t>::max();
  T location = T(std::numeric_limits<typename T::coord_type>::max());
  T velocity = T(std::numeric_limits<typename T::coord_type>::max());
  size_t cellIndex = std::numeric_limits<size_t>::max();
  dt_t timeLeft = std::numeric_limits<dt_t>::max();
  double frequency = std::numeric_limits<double>::max();
  double weight = std::numeric_limits<double>::max();
  double initialWeight = std::numeric_limits<double>::max();
  size_t steps = 0;
  bool on_track = false;
  bool sent = false;

  explicit Particle(size_t id_ = std::numeric_limits<size_t>::max(),
                    const T &location_ = T(std::numeric_limits<double>::max()),
                    const T &velocity_ = T(std::numeric_limits<double>::max()),
                    dt_t timeLeft_ = dt_t(std::numeric_limits<double>::max()))
      : id(id_), location(location_), velocity(velocity_), cellIndex(std::numeric_limits<size_t>::max()),
        timeLeft(timeLeft_), frequency(std::numeric_limits<double>::max()), weight(0), initialWeight(0),
        steps(0), on_track(false) {}

  friend std::ostream &operator<<(std::ostream &os, const Particle &p) {
    return os << "Particle(ID " << p.id << ", location " << p.location << " in cell " << p.cellIndex
              << ", velocity " << p.velocity << ", time " << p.timeLeft << ", steps " << p.steps << ")";
  }
  bool operator==(const Particle &o) const { return id == o.id; }
};

}  // namespace STORM
