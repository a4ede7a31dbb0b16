// ECLIPSE input for the drop-in set-up, over the C-ABI of eclipse_grid.h.
//
// The reference's set-up reads a deck with opm-parser
// (setupGridAndProps.hpp:85-88: Opm::ParseContext, Opm::Parser::parseFile,
// Opm::DeckConstPtr) and hands it to Dune::CpGrid::processEclipseFormat and
// ReservoirPropertyCommon::init.  Those names are kept here for the keywords
// the set-up reads (SPECGRID / DIMENS, COORD, ZCORN, ACTNUM, PERM*, PORO,
// METRIC / FIELD); the reader behind them is mhs_deck_* of libmhs.
#ifndef OPM_PORSOL_AMD_ECLIPSEDECK_HPP
#define OPM_PORSOL_AMD_ECLIPSEDECK_HPP

#include <array>
#include <memory>
#include <stdexcept>
#include <string>

#include <eclipse_grid.h>

namespace Opm {

/// A parsed deck.  Reading stops at the first SOLUTION, SUMMARY, SCHEDULE or
/// REGIONS section; unreadable or unsupported input throws std::runtime_error
/// with the reader's message.
class Deck {
public:
    static std::shared_ptr<const Deck> fromFile(const std::string &path) {
        std::shared_ptr<Deck> d(new Deck());
        d->check(mhs_deck_read_file(&d->h_, path.c_str()), "reading " + path);
        return d;
    }
    static std::shared_ptr<const Deck> fromString(const std::string &text) {
        std::shared_ptr<Deck> d(new Deck());
        d->check(mhs_deck_read_text(&d->h_, text.c_str()), "reading deck text");
        return d;
    }
    ~Deck() {
        if (h_) mhs_deck_destroy(h_);
    }
    Deck(const Deck &) = delete;
    Deck &operator=(const Deck &) = delete;

    /// Logical dimensions (SPECGRID or DIMENS).
    std::array<int, 3> dims() const {
        std::array<int, 3> n{{0, 0, 0}};
        check(mhs_deck_dims(h_, n.data()), "reading the grid dimensions");
        return n;
    }
    /// The C-ABI handle (used by CornerPointGrid and ReservoirPropertyCommon).
    /// The reader keeps no state between calls that a caller could observe,
    /// so const use from several objects in one thread is safe.
    mhs_deck *handle() const { return h_; }
    void check(int rc, const std::string &what) const {
        if (rc != 0)
            throw std::runtime_error("ECLIPSE deck: " + what + ": " +
                                     (h_ ? mhs_deck_last_error(h_) : "out of memory"));
    }

private:
    Deck() = default;
    mhs_deck *h_ = nullptr;
};

typedef std::shared_ptr<const Deck> DeckConstPtr;

/// Stands in for opm-parser's error-handling policy object; the reader here
/// treats every problem as an error.
struct ParseContext {};

/// Opm::Parser of the reference's set-up (setupGridAndProps.hpp:86-88).
class Parser {
public:
    DeckConstPtr parseFile(const std::string &path, const ParseContext & = ParseContext()) const {
        return Deck::fromFile(path);
    }
    DeckConstPtr parseString(const std::string &text, const ParseContext & = ParseContext()) const {
        return Deck::fromString(text);
    }
};

typedef std::shared_ptr<Parser> ParserPtr;

}  // namespace Opm

#endif
