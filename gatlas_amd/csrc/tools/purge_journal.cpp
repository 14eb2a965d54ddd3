// purge_journal: compacts a journal file (purgeJournal.cpp:52-75 semantics):
// keeps the records of bad keys and, unless -d, the RUN_OK state and trial
// times of good keys.  Prints the number of bad keys.
#include <unistd.h>

#include <iostream>
#include <string>

#include "../host/bench.hpp"

int main(int argc, char* argv[]) {
    std::string file;
    bool deleteTimes = false;
    int opt;
    while ((opt = getopt(argc, argv, "hdj:")) != -1) {
        switch (opt) {
            case 'h':
                std::cerr << "usage: " << argv[0] << " -j journalFile [-d]\n"
                          << "\t-d delete benchmark times of good kernels\n";
                return 1;
            case 'j': file = optarg; break;
            case 'd': deleteTimes = true; break;
        }
    }
    if (file.empty()) {
        std::cerr << "error: journal file must be specified" << std::endl;
        return 1;
    }
    gat::Journal j(file);
    if (!j.loadMemo()) {
        std::cerr << "error: could not read journal " << file << std::endl;
        return 1;
    }
    const size_t good = j.memoGood();
    const int bad = j.purgeMemo(deleteTimes);
    if (bad < 0) {
        std::cerr << "error: could not write journal " << file << std::endl;
        return 1;
    }
    std::cout << "journal file " << file << " has " << bad << " bad keys, "
              << (deleteTimes ? "deleted " : "preserved ") << good << " good keys" << std::endl;
    return 0;
}
