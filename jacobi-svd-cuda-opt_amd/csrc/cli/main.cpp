// Entry point of the pipeline CLI (reference src/interfaces/main.cu:17-81).
#include "cli.hpp"

int main(int argc, char *argv[])
{
    return jacobi::svd::cli::run_main(argc, argv);
}
