// Error plumbing shared by every C ABI translation unit.
#pragma once

#include <exception>
#include <string>

namespace mchip {
void set_error(const std::string& s);
}

#define MCHIP_TRY try {
#define MCHIP_CATCH                          \
  }                                          \
  catch (const std::exception& e)            \
  {                                          \
    mchip::set_error(e.what());              \
    return MCHIP_ERR;                        \
  }                                          \
  catch (...)                                \
  {                                          \
    mchip::set_error("unknown C++ exception"); \
    return MCHIP_ERR;                        \
  }                                          \
  return MCHIP_OK;
