// SOPR entry point (apps/SOPR.cu): options, graph load, engine run, metrics report.
#include "app.hpp"

int main(int argc, const char** argv) { return gwh::app_main(argc, argv, gwh::App::SOPR); }
