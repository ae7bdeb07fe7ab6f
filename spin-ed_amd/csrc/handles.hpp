// handles.hpp -- the opaque C handles of include/lattice_symmetries.h: each
// holds a shared_ptr to the engine object (ForeignPtr-style ownership).
#pragma once

#include "engine.hpp"

struct ls_symmetry {
  std::shared_ptr<const spined::Symmetry> p;
};
struct ls_group {
  std::shared_ptr<const spined::Group> p;
};
struct ls_spin_basis {
  std::shared_ptr<spined::Basis> p;
};
struct ls_states {
  std::shared_ptr<spined::HostStates> data;
};
struct ls_interaction {
  std::shared_ptr<const spined::Interaction> p;
};
struct ls_operator {
  std::shared_ptr<spined::Operator> p;
};
