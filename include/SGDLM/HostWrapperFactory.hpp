/*
 * HostWrapperFactory.hpp -- factory of the C++ plugin interface, with the
 * overloads of rSGDLM/src/SGDLM/HostWrapperFactory.hpp:15-24.  Only
 * DOUBLE = double is provided (the reference instantiates DOUBLETYPE=double,
 * rSGDLM/src/Makevars.in:1).  no_gpus is accepted and clamped to one GPU.
 */
#ifndef SGDLM_HOSTWRAPPERFACTORY_HPP_
#define SGDLM_HOSTWRAPPERFACTORY_HPP_

#include "SGDLM/HostWrapper.hpp"

namespace SGDLM {

template <typename DOUBLE> class HostWrapperFactory {
public:
  static HostWrapper<DOUBLE> *create(std::size_t no_gpus);
  static HostWrapper<DOUBLE> *create(std::size_t no_gpus, std::size_t m, std::size_t max_p);
  static HostWrapper<DOUBLE> *create(std::size_t no_gpus, std::size_t m, std::size_t max_p,
                                     bool use_state_evolution_matrix);
};

} // namespace SGDLM

#endif
