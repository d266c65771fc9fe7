/**
 * @file formats.hxx
 * @brief Sparse and dense containers (reference include/loops/container/
 * formats.hxx).
 */
#pragma once

#include <loops/container/coo.hxx>
#include <loops/container/coordinate.hxx>
#include <loops/container/csc.hxx>
#include <loops/container/csr.hxx>
#include <loops/container/vector.hxx>
