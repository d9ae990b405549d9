// Model ingest and solution export (restates include/indus/io.hpp:8-19 of the reference).
// read_mps follows src/io/mps_reader.cpp:48-476 line for line in its observable behaviour,
// including its quirks, so the same file gives the same Model:
//   * fixed and free MPS are both read by whitespace tokenisation (names without blanks);
//   * the objective row's RHS entry is ignored (objective_offset stays 0);
//   * OBJSENSE is read from the line after the header ("OBJSENSE MAX" on one line is minimise);
//   * a column-1 line that is not a known header is parsed as data of the current section;
//   * UP with a negative value and no LO makes the lower bound -1e20;
//   * an E row with range r: r > 0 -> [b, b + r], otherwise [b + r, b];
//   * unparseable numbers are skipped; "D" exponents (1.5D+02) are accepted.
#pragma once
#include <string>

#include "indus/model.hpp"

namespace indus::io {

Model read_mps(const std::string& filepath);

// CPLEX-style LP files, with the behaviour of src/io/lp_reader.cpp:223-793 of the reference:
// MINIMIZE/MAXIMIZE, SUBJECT TO, BOUNDS, GENERALS, BINARY, END sections; "\" comments;
// "name:" labels (the objective's becomes the model name); ranged rows "lhs op expr op rhs";
// bounds "x >= a", "x <= b", "x = v", "x free", "a <= x <= b" with INF/INFINITY; default
// bounds 0 <= x < +inf; a quadratic objective part "[ ... ] / 2" is read into Q when the
// bracket starts a term (after "+ [" the bracket is skipped and its contents are read as
// linear terms, as the reference does).
Model read_lp(const std::string& filepath);

// .sol text (writer.cpp:85-166 of the reference): header block, then one line per column
// and per row with name, value and dual.
void write_solution(const Solution& solution, const Model& model, const std::string& filepath);

// JSON telemetry (writer.cpp:168-264 of the reference).
void write_json(const Solution& solution, const Model& model, const std::string& filepath);

}  // namespace indus::io
