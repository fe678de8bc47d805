// gsi [script] [device]: runs a GS input script (src/main.cu:8-45 of General-Simulator).
#include <fstream>
#include <iostream>
#include <sstream>

#include "gs.h"

int main(int argc, char** argv) {
  std::string file = argc > 1 ? argv[1] : "in.gs";
  gs::log_open(file + ".log");
  if (argc == 1) gs::log_line("Since no input script assigned, default \"in.gs\" is used.");
  else gs::log_line("Input script \"" + file + "\" is used.");
  gs::GS engine;
  if (argc >= 3) {
    engine.Vars["device"] = argv[2];
    gs::log_line(std::string("Gpu device set to ") + argv[2]);
  }
  std::ifstream in(file);
  if (!in) {
    gs::log_line("Input script " + file + " is not found");
    return 1;
  }
  std::stringstream buf;
  buf << in.rdbuf();
  gs::Interpreter interp(engine);
  int rc = interp.run(buf.str());
  return (rc == gs::Code_NORMAL || rc == gs::Code_QUIT) ? 0 : 1;
}
